// Host symbolic phase of the sparse LU / LDL^T solver: adjacency graph,
// multiple-minimum-degree ordering, elimination tree and its postorders,
// column counts, supernode partition and supernodal index sets.
//
// Every routine here restates the step sequence of the corresponding
// Sparspak.jl routine so that the outputs (perm/invp, parent, colcnt,
// xsuper/snode, xlindx/lindx, xlnz/xunz) are identical integer for integer:
// the numeric phase, the storage layout and the relative-index maps uploaded
// to the GPU all derive from them, and minimum-degree tie-breaking depends on
// the exact order of every list operation.
//
// Conventions (same as the reference): all arrays are caller-owned int64,
// index VALUES are 1-based; a pointer p to a length-m array is addressed as
// p[1..m] through a shifted base (B(p)[k] == p[k-1]).
#include <cstdint>
#include <cstring>
#include <cmath>
#include <limits>
#include <vector>
#include <algorithm>

#include "sparspak_hip.h"

namespace {

using i64 = int64_t;

// 1-based view of a 0-based array.
template <typename T> inline T* B(T* p) { return p - 1; }

const i64 kMaxInt = std::numeric_limits<i64>::max();

// ---------------------------------------------------------------------------
// Multiple minimum degree (Liu's genmmd as restated in SpkMMD.jl).
// ---------------------------------------------------------------------------
struct Mmd {
  i64 n;
  const i64* xadj;          // 1-based
  std::vector<i64> adjncy;  // 1-based storage (index 0 unused)
  std::vector<i64> deghead; // 0..n-1
  std::vector<i64> marker, mergeparent, needsupdate, degnext, degprev,
      supersize, elimnext;  // 1-based, size n+1
  i64* invp;                // 1-based

  i64& A(i64 k) { return adjncy[(size_t)k]; }

  void reset_marker() {
    for (i64 k = 1; k <= n; ++k)
      if (marker[k] < kMaxInt) marker[k] = 0;
  }

  // SpkMMD.jl:239-353 (_mmdelim)
  void elim(i64 mdnode, i64 tag) {
    marker[mdnode] = tag;
    i64 elmnt = 0;
    i64 rloc = xadj[mdnode], rlmt = xadj[mdnode + 1] - 1;
    const i64 ilo = xadj[mdnode], ihi = xadj[mdnode + 1] - 1;
    for (i64 i = ilo; i <= ihi; ++i) {
      i64 neighbor = A(i);
      if (neighbor == 0) break;
      if (marker[neighbor] < tag) {
        marker[neighbor] = tag;
        if (invp[neighbor] == 0) {
          A(rloc) = neighbor;
          rloc += 1;
        } else {
          elimnext[neighbor] = elmnt;
          elmnt = neighbor;
        }
      }
    }
    // Merge with reachable nodes from generalized elements.
    while (elmnt > 0) {
      A(rlmt) = -elmnt;
      i64 j = xadj[elmnt], jstop = xadj[elmnt + 1];
      i64 node = A(j);
      while (node != 0) {
        if (node < 0) {
          j = xadj[-node];
          jstop = xadj[-node + 1];
        } else {
          if (marker[node] < tag && degnext[node] >= 0) {
            marker[node] = tag;
            while (rloc >= rlmt) {
              i64 link = -A(rlmt);
              rloc = xadj[link];
              rlmt = xadj[link + 1] - 1;
            }
            A(rloc) = node;
            rloc += 1;
          }
          j += 1;
        }
        if (j >= jstop) break;
        node = A(j);
      }
      elmnt = elimnext[elmnt];
    }
    if (rloc <= rlmt) A(rloc) = 0;
    // For each node in the reachable set ...
    i64 i = xadj[mdnode], istop = xadj[mdnode + 1];
    i64 rnode = A(i);
    while (rnode != 0) {
      if (rnode < 0) {
        i = xadj[-rnode];
        istop = xadj[-rnode + 1];
      } else {
        i64 pvnode = degprev[rnode];
        if (pvnode != 0) {
          i64 nxnode = degnext[rnode];
          if (nxnode > 0) degprev[nxnode] = pvnode;
          if (pvnode > 0)
            degnext[pvnode] = nxnode;
          else
            deghead[(size_t)(-pvnode)] = nxnode;
        }
        // Purge inactive quotient neighbors of rnode.
        i64 xqnbr = xadj[rnode];
        const i64 jlo = xadj[rnode], jhi = xadj[rnode + 1] - 1;
        for (i64 j = jlo; j <= jhi; ++j) {
          i64 neighbor = A(j);
          if (neighbor == 0) break;
          if (marker[neighbor] < tag) {
            A(xqnbr) = neighbor;
            xqnbr += 1;
          }
        }
        i64 nqnbrs = xqnbr - xadj[rnode];
        if (nqnbrs <= 0) {
          supersize[mdnode] += supersize[rnode];
          supersize[rnode] = 0;
          mergeparent[rnode] = mdnode;
          marker[rnode] = kMaxInt;
        } else {
          needsupdate[rnode] = nqnbrs + 1;
          A(xqnbr) = mdnode;
          xqnbr += 1;
          if (xqnbr < xadj[rnode + 1]) A(xqnbr) = 0;
        }
        degprev[rnode] = 0;
        i += 1;
      }
      if (i >= istop) break;
      rnode = A(i);
    }
  }

  // SpkMMD.jl:562-569 (__updateexternaldegree)
  void update_external_degree(i64& deg, i64& mindeg, i64 enode) {
    deg = deg - supersize[enode];
    i64 firstnode = deghead[(size_t)deg];
    deghead[(size_t)deg] = enode;
    degnext[enode] = firstnode;
    degprev[enode] = -deg;
    needsupdate[enode] = 0;
    if (firstnode > 0) degprev[firstnode] = enode;
    if (deg < mindeg) mindeg = deg;
  }

  // SpkMMD.jl:384-559 (_mmdupdate)
  void update(i64 elimhead, i64 delta, i64& mindeg, i64& tag) {
    i64 mindeglimit = mindeg + delta;
    i64 deg = 0, enode = 0;
    i64 elimnode = elimhead;
    while (elimnode > 0) {
      i64 mtag = tag + mindeglimit;
      if (mtag >= kMaxInt) {
        tag = 1;
        mtag = tag + mindeglimit;
        reset_marker();
      }
      i64 q2head = 0, qxhead = 0, elimsize = 0;
      i64 i = xadj[elimnode], istop = xadj[elimnode + 1];
      enode = A(i);
      while (enode != 0) {
        if (enode < 0) {
          i = xadj[-enode];
          istop = xadj[-enode + 1];
        } else {
          if (supersize[enode] != 0) {
            elimsize += supersize[enode];
            marker[enode] = mtag;
            if (needsupdate[enode] > 0) {
              if (needsupdate[enode] != 2) {
                elimnext[enode] = qxhead;
                qxhead = enode;
              } else {
                elimnext[enode] = q2head;
                q2head = enode;
              }
            }
          }
          i += 1;
        }
        if (i >= istop) break;
        enode = A(i);
      }
      // Nodes with two neighbours.
      enode = q2head;
      while (enode > 0) {
        if (needsupdate[enode] > 0) {
          tag += 1;
          deg = elimsize;
          i64 istart = xadj[enode];
          i64 neighbor = A(istart);
          if (neighbor == elimnode) neighbor = A(istart + 1);
          if (invp[neighbor] == 0) {
            deg += supersize[neighbor];
          } else {
            i64 ii = xadj[neighbor], iistop = xadj[neighbor + 1];
            i64 node = A(ii);
            while (node != 0) {
              if (node < 0) {
                ii = xadj[-node];
                iistop = xadj[-node + 1];
              } else {
                if (node != enode && supersize[node] != 0) {
                  if (marker[node] < tag) {
                    marker[node] = tag;
                    deg += supersize[node];
                  } else if (needsupdate[node] > 0) {
                    if (needsupdate[node] == 2) {
                      supersize[enode] += supersize[node];
                      supersize[node] = 0;
                      marker[node] = kMaxInt;
                      mergeparent[node] = enode;
                    }
                    needsupdate[node] = 0;
                    degprev[node] = 0;
                  }
                }
                ii += 1;
              }
              if (ii >= iistop) break;
              node = A(ii);
            }
          }
          update_external_degree(deg, mindeg, enode);
        }
        enode = elimnext[enode];
      }
      // Nodes with more than two neighbours.
      enode = qxhead;
      while (enode > 0) {
        if (needsupdate[enode] > 0) {
          tag += 1;
          deg = elimsize;
          const i64 ilo = xadj[enode], ihi = xadj[enode + 1] - 1;
          for (i64 ii = ilo; ii <= ihi; ++ii) {
            i64 neighbor = A(ii);
            if (neighbor == 0) break;
            if (marker[neighbor] < tag) {
              marker[neighbor] = tag;
              if (invp[neighbor] == 0) {
                deg += supersize[neighbor];
              } else {
                i64 j = xadj[neighbor], jstop = xadj[neighbor + 1];
                i64 node = A(j);
                while (node != 0) {
                  if (node < 0) {
                    j = xadj[-node];
                    jstop = xadj[-node + 1];
                  } else {
                    if (marker[node] < tag) {
                      marker[node] = tag;
                      deg += supersize[node];
                    }
                    j += 1;
                  }
                  if (j >= jstop) break;
                  node = A(j);
                }
              }
            }
          }
          update_external_degree(deg, mindeg, enode);
        }
        enode = elimnext[enode];
      }
      tag = mtag;
      elimnode = elimnext[elimnode];
    }
  }

  // SpkMMD.jl:588-633 (_mmdnumber)
  void number(i64* perm) {
    std::vector<i64> mergelastnum((size_t)n + 1, 0);
    for (i64 k = 1; k <= n; ++k)
      if (mergeparent[k] == 0) mergelastnum[k] = invp[k];
    for (i64 node = 1; node <= n; ++node) {
      i64 parent = mergeparent[node];
      if (parent > 0) {
        i64 root = 0;
        while (parent > 0) {
          root = parent;
          parent = mergeparent[parent];
        }
        i64 num = mergelastnum[root] + 1;
        invp[node] = num;
        mergelastnum[root] = num;
        i64 node2 = node;
        while (node2 != root) {
          parent = mergeparent[node2];
          mergeparent[node2] = root;
          node2 = parent;
        }
      }
    }
    for (i64 k = 1; k <= n; ++k) perm[invp[k]] = k;
  }
};

}  // namespace

extern "C" {

// SpkMMD.jl:78-209 (_generalmmd); delta = 0 as at SpkMMD.jl:79.
int64_t spk_mmd(int64_t n, const int64_t* xadj0, const int64_t* adj0,
                int64_t* perm0, int64_t* invp0) {
  if (n <= 0) return -1;
  Mmd m;
  m.n = n;
  m.xadj = B(xadj0);
  const i64 nadj = m.xadj[n + 1] - 1;
  // One spare slot past the end: the reference never reads it, but a read of
  // adjncy[xadj[n+1]] then stays inside the allocation.
  m.adjncy.assign((size_t)nadj + 2, 0);
  for (i64 k = 1; k <= nadj; ++k) m.adjncy[(size_t)k] = adj0[k - 1];
  m.deghead.assign((size_t)n, 0);
  m.marker.assign((size_t)n + 1, 0);
  m.mergeparent.assign((size_t)n + 1, 0);
  m.needsupdate.assign((size_t)n + 1, 0);
  m.degnext.assign((size_t)n + 1, 0);
  m.degprev.assign((size_t)n + 1, 0);
  m.supersize.assign((size_t)n + 1, 1);
  m.elimnext.assign((size_t)n + 1, 0);
  i64* perm = B(perm0);
  i64* invp = B(invp0);
  m.invp = invp;
  const i64 delta = 0;
  const i64* xadj = m.xadj;

  for (i64 node = 1; node <= n; ++node) {
    i64 ndeg = xadj[node + 1] - xadj[node];
    if (ndeg < 0 || ndeg > n - 1) return -2;
    i64 fnode = m.deghead[(size_t)ndeg];
    m.deghead[(size_t)ndeg] = node;
    m.degnext[node] = fnode;
    if (fnode > 0) m.degprev[fnode] = node;
    m.degprev[node] = -ndeg;
  }
  i64 num = 1;
  for (i64 k = 1; k <= n; ++k) invp[k] = 0;
  i64 mdnode = m.deghead[0];
  while (mdnode > 0) {
    m.marker[mdnode] = kMaxInt;
    invp[mdnode] = num;
    num += 1;
    mdnode = m.degnext[mdnode];
  }
  m.deghead[0] = 0;
  i64 tag = 1;
  i64 mindeg = 1;
  while (num <= n) {
    while (m.deghead[(size_t)mindeg] <= 0) {
      mindeg += 1;
      if (mindeg >= n) return -3;
    }
    i64 mindeglimit = mindeg + delta;
    i64 elimhead = 0;
    bool done = false;
    for (;;) {
      mdnode = m.deghead[(size_t)mindeg];
      bool pass = false;
      while (mdnode <= 0) {
        mindeg += 1;
        if (mindeg > mindeglimit) {
          pass = true;
          break;
        }
        mdnode = m.deghead[(size_t)mindeg];
      }
      if (pass) break;
      i64 mdnextnode = m.degnext[mdnode];
      m.deghead[(size_t)mindeg] = mdnextnode;
      if (mdnextnode > 0) m.degprev[mdnextnode] = -mindeg;
      invp[mdnode] = num;
      if (num + m.supersize[mdnode] > n) {
        done = true;
        break;
      }
      tag += 1;
      if (tag >= kMaxInt) {
        tag = 1;
        m.reset_marker();
      }
      m.elim(mdnode, tag);
      num += m.supersize[mdnode];
      m.elimnext[mdnode] = elimhead;
      elimhead = mdnode;
    }
    if (done) break;
    if (num > n) break;
    m.update(elimhead, delta, mindeg, tag);
  }
  m.number(perm);
  return 0;
}

// SpkGraph.jl:298-316 (isstructuresymmetric)
int64_t spk_isstructuresymmetric(int64_t nv, const int64_t* xadj0,
                                 const int64_t* adj0) {
  const i64* xadj = B(xadj0);
  const i64* adj = B(adj0);
  for (i64 i = 1; i <= nv; ++i) {
    for (i64 k = xadj[i]; k <= xadj[i + 1] - 1; ++k) {
      i64 j = adj[k];
      bool found = false;
      for (i64 kk = xadj[j]; kk <= xadj[j + 1] - 1; ++kk)
        if (adj[kk] == i) {
          found = true;
          break;
        }
      if (!found) return 0;
    }
  }
  return 1;
}

// SpkGraph.jl:94-268 (makestructuresymmetric, counting pass) -> the edge
// count of the symmetrised graph, or nedges itself when already symmetric.
int64_t spk_symmetrized_nedges(int64_t nv, const int64_t* xadj0,
                               const int64_t* adj0) {
  if (spk_isstructuresymmetric(nv, xadj0, adj0)) return B(xadj0)[nv + 1] - 1;
  const i64* xadj = B(xadj0);
  const i64* adj = B(adj0);
  i64 nedges = xadj[nv + 1] - 1;
  std::vector<i64> first((size_t)nv + 1);
  for (i64 i = 1; i <= nv; ++i) first[i] = xadj[i];
  for (i64 i = 1; i <= nv; ++i) {
    for (i64 j = first[i]; j <= xadj[i + 1] - 1; ++j) {
      i64 k = adj[j];
      if (k < i) {
        nedges += 1;
      } else {
        while (first[k] <= xadj[k + 1]) {
          if (first[k] == xadj[k + 1]) {
            nedges += 1;
            break;
          }
          i64 m = adj[first[k]];
          if (m < i) {
            nedges += 1;
            first[k] += 1;
          } else if (m == i) {
            first[k] += 1;
            break;
          } else {
            nedges += 1;
            break;
          }
        }
      }
    }
  }
  return nedges;
}

// SpkGraph.jl:94-289 (makestructuresymmetric + sortgraph!). Writes the
// symmetrised graph into (hxadj, hadj); hadj must hold
// spk_symmetrized_nedges() entries. A graph that is already symmetric is
// copied unchanged (no sort), as the reference returns early.
int64_t spk_makestructuresymmetric(int64_t nv, const int64_t* xadj0,
                                   const int64_t* adj0, int64_t* hxadj0,
                                   int64_t* hadj0) {
  const i64* xadj = B(xadj0);
  const i64* adj = B(adj0);
  i64* hx = B(hxadj0);
  i64* ha = B(hadj0);
  if (spk_isstructuresymmetric(nv, xadj0, adj0)) {
    for (i64 i = 1; i <= nv + 1; ++i) hx[i] = xadj[i];
    for (i64 k = 1; k <= xadj[nv + 1] - 1; ++k) ha[k] = adj[k];
    return 1;
  }
  std::vector<i64> c((size_t)nv + 1, 0), first((size_t)nv + 1);
  for (i64 i = 1; i <= nv; ++i) first[i] = xadj[i];
  // counting pass
  for (i64 i = 1; i <= nv; ++i) {
    for (i64 j = first[i]; j <= xadj[i + 1] - 1; ++j) {
      i64 k = adj[j];
      c[i] += 1;
      if (k < i) {
        c[k] += 1;
      } else {
        while (first[k] <= xadj[k + 1]) {
          if (first[k] == xadj[k + 1]) {
            c[k] += 1;
            break;
          }
          i64 m = adj[first[k]];
          if (m < i) {
            c[m] += 1;
            c[k] += 1;
            first[k] += 1;
          } else if (m == i) {
            c[k] += 1;
            first[k] += 1;
            break;
          } else {
            c[k] += 1;
            break;
          }
        }
      }
    }
  }
  hx[1] = 1;
  for (i64 i = 1; i <= nv; ++i) {
    first[i] = xadj[i];
    hx[i + 1] = hx[i] + c[i];
    c[i] = hx[i];
  }
  const i64 hn = hx[nv + 1] - 1;
  std::vector<i64> h((size_t)hn + 1, 0);
  // recording pass
  for (i64 i = 1; i <= nv; ++i) {
    for (i64 j = first[i]; j <= xadj[i + 1] - 1; ++j) {
      i64 k = adj[j];
      h[c[i]] = k;
      c[i] += 1;
      if (k < i) {
        h[c[k]] = i;
        c[k] += 1;
      } else {
        while (first[k] <= xadj[k + 1]) {
          if (first[k] == xadj[k + 1]) {
            h[c[k]] = i;
            c[k] += 1;
            break;
          }
          i64 m = adj[first[k]];
          if (m < i) {
            h[c[m]] = k;
            c[m] += 1;
            h[c[k]] = m;
            c[k] += 1;
            first[k] += 1;
          } else if (m == i) {
            h[c[k]] = m;
            c[k] += 1;
            first[k] += 1;
            break;
          } else {
            h[c[k]] = i;
            c[k] += 1;
            break;
          }
        }
      }
    }
  }
  // sortgraph!: transpose of the (symmetric) graph lists each vertex's
  // neighbours in increasing order.
  std::vector<i64> pos((size_t)nv + 2);
  for (i64 i = 1; i <= nv + 1; ++i) pos[i] = hx[i];
  for (i64 i = 1; i <= nv; ++i)
    for (i64 j = hx[i]; j <= hx[i + 1] - 1; ++j) {
      i64 k = h[j];
      ha[pos[k]] = i;
      pos[k] += 1;
    }
  return 0;
}

// SpkETree.jl:62-88 (_findetree!)
void spk_etree(int64_t n, const int64_t* xadj0, const int64_t* adj0,
               const int64_t* perm0, const int64_t* invp0, int64_t* parent0) {
  const i64* xadj = B(xadj0);
  const i64* adj = B(adj0);
  const i64* rperm = B(perm0);
  const i64* rinvp = B(invp0);
  i64* parent = B(parent0);
  std::vector<i64> ancestor((size_t)n + 1, 0);
  for (i64 i = 1; i <= n; ++i) {
    parent[i] = 0;
    ancestor[i] = 0;
    i64 vertex = rperm[i];
    for (i64 j = xadj[vertex]; j <= xadj[vertex + 1] - 1; ++j) {
      i64 nbr = rinvp[adj[j]];
      if (nbr < i) {
        while (ancestor[nbr] != 0 && ancestor[nbr] != i) {
          i64 next = ancestor[nbr];
          ancestor[nbr] = i;
          nbr = next;
        }
        if (ancestor[nbr] == 0) {
          parent[nbr] = i;
          ancestor[nbr] = i;
        }
      }
    }
  }
}

}  // extern "C"

namespace {

// SpkETree.jl:153-174 (_binarytree!)
void binarytree(i64 n, const i64* parent, i64* fson, i64* brother) {
  for (i64 k = 1; k <= n; ++k) fson[k] = brother[k] = 0;
  i64 lroot = n;
  for (i64 vertex = n - 1; vertex >= 1; --vertex) {
    i64 vpar = parent[vertex];
    if (vpar <= 0 || vpar == vertex) {
      brother[lroot] = vertex;
      lroot = vertex;
    } else {
      brother[vertex] = fson[vpar];
      fson[vpar] = vertex;
    }
  }
  brother[lroot] = 0;
}

// SpkETree.jl:239-274 (_weightedbinarytree!). A child whose weight is at
// least the current last son's goes to the front of the sibling list,
// otherwise it becomes the new last son.
void weightedbinarytree(i64 n, const i64* parent, const i64* weight, i64* fson,
                        i64* brother) {
  std::vector<i64> lson((size_t)n + 1, 0);
  for (i64 k = 1; k <= n; ++k) fson[k] = brother[k] = 0;
  i64 lroot = n;
  for (i64 vertex = n - 1; vertex >= 1; --vertex) {
    i64 vpar = parent[vertex];
    if (vpar <= 0 || vpar == vertex) {
      brother[lroot] = vertex;
      lroot = vertex;
    } else {
      i64 vlastson = lson[vpar];
      if (vlastson != 0) {
        if (weight[vertex] >= weight[vlastson]) {
          brother[vertex] = fson[vpar];
          fson[vpar] = vertex;
        } else {
          brother[vlastson] = vertex;
          lson[vpar] = vertex;
        }
      } else {
        fson[vpar] = vertex;
        lson[vpar] = vertex;
      }
    }
  }
  brother[lroot] = 0;
}

// SpkETree.jl:190-220 (_postordertree!)
void postordertree(i64 root, i64* fson, i64* brother, i64* rinvpos,
                   i64* parent, i64* stack) {
  i64 num = 0, top = 0, vertex = root;
  while (vertex > 0) {
    while (vertex > 0) {
      top += 1;
      stack[top] = vertex;
      vertex = fson[vertex];
    }
    while (vertex <= 0 && top > 0) {
      vertex = stack[top];
      top -= 1;
      num += 1;
      rinvpos[vertex] = num;
      vertex = brother[vertex];
    }
  }
  for (i64 v = 1; v <= num; ++v) {
    i64 nuvertex = rinvpos[v];
    i64 vpar = parent[v];
    if (vpar > 0) vpar = rinvpos[vpar];
    brother[nuvertex] = vpar;
  }
  for (i64 v = 1; v <= root; ++v) parent[v] = brother[v];
}

}  // namespace

extern "C" {

// SpkETree.jl:106-139 (_getpostorder!, with or without weight). Updates
// parent, perm/invp and (when given) permutes weight into the new order.
void spk_postorder(int64_t n, int64_t* parent0, int64_t* perm0, int64_t* invp0,
                   int64_t* weight0) {
  i64* parent = B(parent0);
  i64* rperm = B(perm0);
  i64* rinvp = B(invp0);
  std::vector<i64> fson((size_t)n + 1), brother((size_t)n + 1),
      stack((size_t)n + 1), invpos((size_t)n + 1, 0);
  if (weight0)
    weightedbinarytree(n, parent, B(weight0), fson.data(), brother.data());
  else
    binarytree(n, parent, fson.data(), brother.data());
  postordertree(n, fson.data(), brother.data(), invpos.data(), parent,
                stack.data());
  if (weight0) {
    i64* weight = B(weight0);
    for (i64 k = 1; k <= n; ++k) stack[invpos[k]] = weight[k];
    for (i64 k = 1; k <= n; ++k) weight[k] = stack[k];
  }
  for (i64 k = 1; k <= n; ++k) stack[k] = invpos[rinvp[k]];
  for (i64 k = 1; k <= n; ++k) rinvp[k] = stack[k];
  for (i64 k = 1; k <= n; ++k) rperm[rinvp[k]] = k;
}

// SpkSymFct.jl:37-59 (_findcolumncounts!). Returns nnz(L) incl. diagonal.
int64_t spk_colcounts(int64_t n, const int64_t* xadj0, const int64_t* adj0,
                      const int64_t* perm0, const int64_t* invp0,
                      const int64_t* parent0, int64_t* colcnt0) {
  const i64* xadj = B(xadj0);
  const i64* adj = B(adj0);
  const i64* perm = B(perm0);
  const i64* invp = B(invp0);
  const i64* parent = B(parent0);
  i64* colcnt = B(colcnt0);
  std::vector<i64> marker((size_t)n + 1, 0);
  i64 nlnz = n;
  for (i64 i = 1; i <= n; ++i) {
    marker[i] = i;
    colcnt[i] = 1;
    for (i64 k = xadj[perm[i]]; k <= xadj[perm[i] + 1] - 1; ++k) {
      i64 j = invp[adj[k]];
      if (j < i) {
        while (marker[j] != i) {
          colcnt[j] += 1;
          nlnz += 1;
          marker[j] = i;
          j = parent[j];
        }
      }
    }
  }
  return nlnz;
}

// SpkSymFct.jl:415-464 (_findsupernodes!). xsuper needs n+1 entries. Returns
// nsuper; *nofsub receives the total index-set length.
int64_t spk_supernodes(int64_t n, const int64_t* parent0,
                       const int64_t* colcnt0, int64_t maxsize,
                       int64_t* xsuper0, int64_t* snode0, int64_t* nofsub) {
  const i64* parent = B(parent0);
  const i64* colcnt = B(colcnt0);
  i64* xsuper = B(xsuper0);
  i64* snode = B(snode0);
  std::vector<i64> marker((size_t)n + 2, 0);
  // Fundamental supernodes.
  i64 nsuper = 1;
  xsuper[1] = 1;
  for (i64 kcol = 2; kcol <= n; ++kcol) {
    if (parent[kcol - 1] == kcol && colcnt[kcol - 1] == colcnt[kcol] + 1)
      continue;
    nsuper += 1;
    xsuper[nsuper] = kcol;
  }
  xsuper[nsuper + 1] = n + 1;
  // Split the wide ones (Float64 arithmetic and floor as in the reference).
  for (i64 jsup = 1; jsup <= nsuper; ++jsup) {
    i64 supsize = xsuper[jsup + 1] - xsuper[jsup];
    if (supsize > maxsize) {
      double delta = (double)supsize / (1.0 + (double)supsize / (double)maxsize);
      i64 limit = (i64)std::floor((double)xsuper[jsup + 1] - delta / 2.0);
      i64 k = 1;
      i64 index = (i64)std::floor((double)xsuper[jsup] + delta);
      while (index < limit) {
        marker[index] = 1;
        k += 1;
        index = (i64)std::floor((double)xsuper[jsup] + (double)k * delta);
      }
    }
  }
  nsuper = 1;
  snode[1] = 1;
  i64 nsub = colcnt[1];
  for (i64 kcol = 2; kcol <= n; ++kcol) {
    if (marker[kcol] != 1) {
      if (parent[kcol - 1] == kcol) {
        if (colcnt[kcol - 1] == colcnt[kcol] + 1) {
          snode[kcol] = nsuper;
          continue;
        }
      }
    }
    nsuper += 1;
    snode[kcol] = nsuper;
    xsuper[nsuper] = kcol;
    nsub += colcnt[kcol];
  }
  xsuper[nsuper + 1] = n + 1;
  if (nofsub) *nofsub = nsub;
  return nsuper;
}

// SpkSparseBase.jl:253-284 (findnonzeroindexs): column pointers of the LU
// panels; every column of supernode J gets colcnt[fj] slots in lnz and
// colcnt[fj] - width slots in unz.
void spk_nonzero_indexs(int64_t n, const int64_t* colcnt0, int64_t nsuper,
                        const int64_t* xsuper0, int64_t* xlnz0,
                        int64_t* xunz0) {
  const i64* colcnt = B(colcnt0);
  const i64* xsuper = B(xsuper0);
  i64* xlnz = B(xlnz0);
  i64* xunz = B(xunz0);
  i64 upoint = 1, point = 1;
  for (i64 ksup = 1; ksup <= nsuper; ++ksup) {
    i64 fstcol = xsuper[ksup], lstcol = xsuper[ksup + 1] - 1;
    i64 width = lstcol - fstcol + 1;
    for (i64 jcol = fstcol; jcol <= lstcol; ++jcol) {
      xlnz[jcol] = point;
      xunz[jcol] = upoint;
      point += colcnt[fstcol];
      upoint += colcnt[fstcol] - width;
    }
  }
  xlnz[n + 1] = point;
  xunz[n + 1] = upoint;
}

// SPD storage pointers (SpkSparseSpdBase.jl:359-375 _findnonzeroindexs!):
// rectangular jlen x nj panels, no U.
void spk_nonzero_indexs_spd(int64_t n, const int64_t* colcnt0, int64_t nsuper,
                            const int64_t* xsuper0, int64_t* xlnz0) {
  const i64* colcnt = B(colcnt0);
  const i64* xsuper = B(xsuper0);
  i64* xlnz = B(xlnz0);
  i64 point = 1;
  for (i64 ksup = 1; ksup <= nsuper; ++ksup) {
    i64 fstcol = xsuper[ksup], lstcol = xsuper[ksup + 1] - 1;
    for (i64 jcol = fstcol; jcol <= lstcol; ++jcol) {
      xlnz[jcol] = point;
      point += colcnt[fstcol];
    }
  }
  xlnz[n + 1] = point;
}

// SpkSymFct.jl:105-246 (_symbolicfact!). Returns 0, or -1 on the
// reference's "Inconsistency in data structure" error.
int64_t spk_symbfact(int64_t n, const int64_t* xadj0, const int64_t* adj0,
                     const int64_t* perm0, const int64_t* invp0,
                     const int64_t* colcnt0, int64_t nsuper,
                     const int64_t* xsuper0, const int64_t* snode0,
                     int64_t nofsub, int64_t* xlindx0, int64_t* lindx0) {
  const i64* xadj = B(xadj0);
  const i64* adj = B(adj0);
  const i64* perm = B(perm0);
  const i64* invp = B(invp0);
  const i64* colcnt = B(colcnt0);
  const i64* xsuper = B(xsuper0);
  const i64* snode = B(snode0);
  i64* xlindx = B(xlindx0);
  i64* lindx = B(lindx0);
  std::vector<i64> marker((size_t)n + 1, 0), mrglnk((size_t)nsuper + 1, 0);
  std::vector<i64> rchlnk((size_t)n + 2, 0);  // indices 0..n (+ tail slot)
  const i64 head = 0, tail = n + 1;
  i64 nzend = 0, point = 1;
  for (i64 ksup = 1; ksup <= nsuper; ++ksup) {
    mrglnk[ksup] = 0;
    i64 fstcol = xsuper[ksup];
    xlindx[ksup] = point;
    point += colcnt[fstcol];
  }
  xlindx[nsuper + 1] = point;
  for (i64 ksup = 1; ksup <= nsuper; ++ksup) {
    i64 fstcol = xsuper[ksup], lstcol = xsuper[ksup + 1] - 1;
    i64 width = lstcol - fstcol + 1, lngth = colcnt[fstcol];
    i64 knz = 0;
    rchlnk[head] = tail;
    i64 jsup = mrglnk[ksup];
    if (jsup > 0) {
      i64 jwidth = xsuper[jsup + 1] - xsuper[jsup];
      i64 jnzbeg = xlindx[jsup] + jwidth, jnzend = xlindx[jsup + 1] - 1;
      for (i64 jptr = jnzend; jptr >= jnzbeg; --jptr) {
        i64 newi = lindx[jptr];
        knz += 1;
        marker[newi] = ksup;
        rchlnk[newi] = rchlnk[head];
        rchlnk[head] = newi;
      }
      jsup = mrglnk[jsup];
      while (jsup != 0 && knz < lngth) {
        jwidth = xsuper[jsup + 1] - xsuper[jsup];
        jnzbeg = xlindx[jsup] + jwidth;
        jnzend = xlindx[jsup + 1] - 1;
        i64 nexti = head;
        for (i64 jptr = jnzbeg; jptr <= jnzend; ++jptr) {
          i64 newi = lindx[jptr];
          i64 i = nexti;
          nexti = rchlnk[i];
          while (newi > nexti) {
            i = nexti;
            nexti = rchlnk[i];
          }
          if (newi < nexti) {
            knz += 1;
            rchlnk[i] = newi;
            rchlnk[newi] = nexti;
            marker[newi] = ksup;
            nexti = newi;
          }
        }
        jsup = mrglnk[jsup];
      }
    }
    if (knz < lngth) {
      i64 node = perm[fstcol];
      for (i64 kptr = xadj[node]; kptr <= xadj[node + 1] - 1; ++kptr) {
        i64 newi = invp[adj[kptr]];
        if (newi > fstcol && marker[newi] != ksup) {
          i64 nexti = head, i = nexti;
          nexti = rchlnk[i];
          while (newi > nexti) {
            i = nexti;
            nexti = rchlnk[i];
          }
          knz += 1;
          rchlnk[i] = newi;
          rchlnk[newi] = nexti;
          marker[newi] = ksup;
        }
      }
    }
    if (rchlnk[head] != fstcol) {
      rchlnk[fstcol] = rchlnk[head];
      rchlnk[head] = fstcol;
      knz += 1;
    }
    i64 nzbeg = nzend + 1;
    nzend = nzend + knz;
    if (nzend + 1 != xlindx[ksup + 1]) return -1;
    if (nzend > nofsub) return -1;
    i64 i = head;
    for (i64 kptr = nzbeg; kptr <= nzend; ++kptr) {
      i = rchlnk[i];
      lindx[kptr] = i;
    }
    if (lngth > width) {
      i64 pcol = lindx[xlindx[ksup] + width];
      i64 psup = snode[pcol];
      mrglnk[ksup] = mrglnk[psup];
      mrglnk[psup] = ksup;
    }
  }
  return 0;
}

// SpkSymFct.jl:372-380 (_findsupernodetree!): supernodal parent vector
// (0 for a root). The reference assumes a single root at nsuper; here every
// supernode whose index list ends with its own columns is a root.
void spk_supernode_tree(int64_t nsuper, const int64_t* xsuper0,
                        const int64_t* snode0, const int64_t* xlindx0,
                        const int64_t* lindx0, int64_t* sparent0) {
  const i64* xsuper = B(xsuper0);
  const i64* snode = B(snode0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  i64* sparent = B(sparent0);
  for (i64 jsup = 1; jsup <= nsuper; ++jsup) {
    i64 j = xlindx[jsup] + xsuper[jsup + 1] - xsuper[jsup];
    sparent[jsup] = (j <= xlindx[jsup + 1] - 1) ? snode[lindx[j]] : 0;
  }
}

}  // extern "C"
