"""ORACLE — test infrastructure only (used by tests/ and tests/golden/).

A literal pure-Python restatement of Sparspak.jl's integer ordering code,
kept apart from the product's C++ restatement (sparspak.jl_amd/csrc/
spk_symbolic.cpp) so that the two can be compared node for node:

  mmd()                SpkMMD.jl:78-209   _generalmmd (delta = 0)
    _mmdelim()         SpkMMD.jl:239-353
    _mmdupdate()       SpkMMD.jl:384-559  (+ __updateexternaldegree :562-569)
    _mmdnumber()       SpkMMD.jl:588-633
  findetree()          SpkETree.jl:62-88
  binarytree()         SpkETree.jl:153-174
  weightedbinarytree() SpkETree.jl:239-274
  postordertree()      SpkETree.jl:190-220
  getpostorder()       SpkETree.jl:106-139

Arrays are 1-based as in the reference (index 0 unused). Every decision that
breaks a tie, merges nodes or orders children is appended to an optional
trace list as a tuple whose second field is the reference line that made it:

  ("isolated", 126, node, number)          isolated node numbered (list order)
  ("select", 157, mindeg, node, candidates) head of the minimum-degree list
                                            taken; candidates = the whole list
  ("mass", 336, node, into)                 reachable node absorbed (no
                                            active neighbour left)
  ("indist", 484, node, into)               indistinguishable node merged
  ("degree", 563, node, external_degree)    degree update
  ("son", 261|264|267, vertex, parent, decision)   weighted binary tree:
                                            "first" (front), "last" (appended)

The tag reset of the reference (when tag reaches typemax(Int64)) cannot
happen at test sizes and is restated only as an assertion.
"""
MAXINT = (1 << 63) - 1


def mmd(n, xadj, adj, trace=None):
    """_generalmmd (SpkMMD.jl:78-209) on a 1-based symmetric adjacency
    structure without diagonal: returns (perm, invp), 1-based lists of n."""
    X = [0] + list(xadj)                      # X[v] = xadj[v], v = 1..n+1
    A = [0] + list(adj[: X[n + 1] - 1])       # :84 copy of the adjacency
    deghead = [0] * (n + 1)                   # :86-87 indices 0..n-1
    marker = [0] * (n + 1)
    mergeparent = [0] * (n + 1)
    needsupdate = [0] * (n + 1)
    degnext = [0] * (n + 1)
    degprev = [0] * (n + 1)
    supersize = [1] * (n + 1)
    elimnext = [0] * (n + 1)
    invp = [0] * (n + 1)
    st = dict(X=X, A=A, deghead=deghead, marker=marker, mergeparent=mergeparent,
              needsupdate=needsupdate, degnext=degnext, degprev=degprev,
              supersize=supersize, elimnext=elimnext, invp=invp, trace=trace)
    for node in range(1, n + 1):              # :108-115 degree lists (LIFO)
        ndeg = X[node + 1] - X[node]
        fnode = deghead[ndeg]
        deghead[ndeg] = node
        degnext[node] = fnode
        if fnode > 0:
            degprev[fnode] = node
        degprev[node] = -ndeg
    num = 1
    mdnode = deghead[0]                       # :124-131 isolated nodes first
    while mdnode > 0:
        marker[mdnode] = MAXINT
        invp[mdnode] = num
        if trace is not None:
            trace.append(("isolated", 126, mdnode, num))
        num += 1
        mdnode = degnext[mdnode]
    deghead[0] = 0
    tag = 1
    mindeg = 1
    done = False
    while num <= n and not done:              # :139
        while deghead[mindeg] <= 0:
            mindeg += 1
        mindeglimit = mindeg                  # delta = 0
        elimhead = 0
        while True:                           # :153
            mdnode = deghead[mindeg]
            while mdnode <= 0:
                mindeg += 1
                if mindeg > mindeglimit:
                    break
                mdnode = deghead[mindeg]
            if mdnode <= 0:                   # @goto pass
                break
            if trace is not None:
                cand, v = [], mdnode
                while v > 0:
                    cand.append(v)
                    v = degnext[v]
                trace.append(("select", 157, mindeg, mdnode, cand))
            mdnextnode = degnext[mdnode]      # :168-172
            deghead[mindeg] = mdnextnode
            if mdnextnode > 0:
                degprev[mdnextnode] = -mindeg
            invp[mdnode] = num
            if num + supersize[mdnode] > n:   # @goto main
                done = True
                break
            tag += 1
            assert tag < MAXINT
            _mmdelim(st, mdnode, tag)
            num += supersize[mdnode]
            elimnext[mdnode] = elimhead
            elimhead = mdnode
        if done:
            break
        if num > n:                           # :201-203
            break
        mindeg, tag = _mmdupdate(st, elimhead, n, mindeg, tag)
    perm = _mmdnumber(n, invp, mergeparent)
    return perm[1:], invp[1:]


def _mmdelim(st, mdnode, tag):
    """SpkMMD.jl:239-353."""
    X, A = st["X"], st["A"]
    marker, invp, elimnext = st["marker"], st["invp"], st["elimnext"]
    degnext, degprev, deghead = st["degnext"], st["degprev"], st["deghead"]
    supersize, mergeparent, needsupdate = st["supersize"], st["mergeparent"], st["needsupdate"]
    trace = st["trace"]
    marker[mdnode] = tag
    elmnt = 0
    rloc = X[mdnode]
    rlmt = X[mdnode + 1] - 1
    for i in range(X[mdnode], X[mdnode + 1]):     # :252-263
        neighbor = A[i]
        if neighbor == 0:
            break
        if marker[neighbor] < tag:
            marker[neighbor] = tag
            if invp[neighbor] == 0:
                A[rloc] = neighbor
                rloc += 1
            else:
                elimnext[neighbor] = elmnt
                elmnt = neighbor
    while elmnt > 0:                              # :267-291
        A[rlmt] = -elmnt
        j = X[elmnt]
        jstop = X[elmnt + 1]
        node = A[j]
        while node != 0:
            if node < 0:
                j = X[-node]
                jstop = X[-node + 1]
            else:
                if marker[node] < tag and degnext[node] >= 0:
                    marker[node] = tag
                    while rloc >= rlmt:
                        link = -A[rlmt]
                        rloc = X[link]
                        rlmt = X[link + 1] - 1
                    A[rloc] = node
                    rloc += 1
                j += 1
            if j >= jstop:
                break
            node = A[j]
        elmnt = elimnext[elmnt]
    if rloc <= rlmt:
        A[rloc] = 0
    i = X[mdnode]                                 # :296-352
    istop = X[mdnode + 1]
    rnode = A[i]
    while rnode != 0:
        if rnode < 0:
            i = X[-rnode]
            istop = X[-rnode + 1]
        else:
            pvnode = degprev[rnode]
            if pvnode != 0:
                nxnode = degnext[rnode]
                if nxnode > 0:
                    degprev[nxnode] = pvnode
                if pvnode > 0:
                    degnext[pvnode] = nxnode
                else:
                    deghead[-pvnode] = nxnode
            xqnbr = X[rnode]
            for j in range(X[rnode], X[rnode + 1]):
                neighbor = A[j]
                if neighbor == 0:
                    break
                if marker[neighbor] < tag:
                    A[xqnbr] = neighbor
                    xqnbr += 1
            nqnbrs = xqnbr - X[rnode]
            if nqnbrs <= 0:
                supersize[mdnode] += supersize[rnode]
                supersize[rnode] = 0
                mergeparent[rnode] = mdnode
                marker[rnode] = MAXINT
                if trace is not None:
                    trace.append(("mass", 336, rnode, mdnode))
            else:
                needsupdate[rnode] = nqnbrs + 1
                A[xqnbr] = mdnode
                xqnbr += 1
                if xqnbr < X[rnode + 1]:
                    A[xqnbr] = 0
            degprev[rnode] = 0
            i += 1
        if i >= istop:
            break
        rnode = A[i]


def _update_external_degree(st, deg, mindeg, enode):
    """__updateexternaldegree, SpkMMD.jl:562-569."""
    supersize, deghead, degnext = st["supersize"], st["deghead"], st["degnext"]
    degprev, needsupdate = st["degprev"], st["needsupdate"]
    deg = deg - supersize[enode]
    firstnode = deghead[deg]
    deghead[deg] = enode
    degnext[enode] = firstnode
    degprev[enode] = -deg
    needsupdate[enode] = 0
    if firstnode > 0:
        degprev[firstnode] = enode
    if deg < mindeg:
        mindeg = deg
    if st["trace"] is not None:
        st["trace"].append(("degree", 563, enode, deg))
    return deg, mindeg


def _mmdupdate(st, elimhead, neqns, mindeg, tag):
    """SpkMMD.jl:384-559 (delta = 0)."""
    X, A = st["X"], st["A"]
    marker, invp, elimnext = st["marker"], st["invp"], st["elimnext"]
    degprev = st["degprev"]
    supersize, mergeparent, needsupdate = st["supersize"], st["mergeparent"], st["needsupdate"]
    trace = st["trace"]
    mindeglimit = mindeg
    elimnode = elimhead
    while elimnode > 0:
        mtag = tag + mindeglimit
        assert mtag < MAXINT
        q2head = qxhead = elimsize = 0
        i = X[elimnode]
        istop = X[elimnode + 1]
        enode = A[i]
        while enode != 0:                         # :413-439
            if enode < 0:
                i = X[-enode]
                istop = X[-enode + 1]
            else:
                if supersize[enode] != 0:
                    elimsize += supersize[enode]
                    marker[enode] = mtag
                    if needsupdate[enode] > 0:
                        if needsupdate[enode] != 2:
                            elimnext[enode] = qxhead
                            qxhead = enode
                        else:
                            elimnext[enode] = q2head
                            q2head = enode
                i += 1
            if i >= istop:
                break
            enode = A[i]
        enode = q2head                            # :443-502
        while enode > 0:
            if needsupdate[enode] > 0:
                tag += 1
                deg = elimsize
                istart = X[enode]
                neighbor = A[istart]
                if neighbor == elimnode:
                    neighbor = A[istart + 1]
                if invp[neighbor] == 0:
                    deg += supersize[neighbor]
                else:
                    i = X[neighbor]
                    istop = X[neighbor + 1]
                    node = A[i]
                    while node != 0:
                        if node < 0:
                            i = X[-node]
                            istop = X[-node + 1]
                        else:
                            if node != enode and supersize[node] != 0:
                                if marker[node] < tag:
                                    marker[node] = tag
                                    deg += supersize[node]
                                elif needsupdate[node] > 0:
                                    if needsupdate[node] == 2:
                                        supersize[enode] += supersize[node]
                                        supersize[node] = 0
                                        marker[node] = MAXINT
                                        mergeparent[node] = enode
                                        if trace is not None:
                                            trace.append(("indist", 484, node, enode))
                                    needsupdate[node] = 0
                                    degprev[node] = 0
                            i += 1
                        if i >= istop:
                            break
                        node = A[i]
                deg, mindeg = _update_external_degree(st, deg, mindeg, enode)
            enode = elimnext[enode]
        enode = qxhead                            # :506-552
        while enode > 0:
            if needsupdate[enode] > 0:
                tag += 1
                deg = elimsize
                for i in range(X[enode], X[enode + 1]):
                    neighbor = A[i]
                    if neighbor == 0:
                        break
                    if marker[neighbor] < tag:
                        marker[neighbor] = tag
                        if invp[neighbor] == 0:
                            deg += supersize[neighbor]
                        else:
                            j = X[neighbor]
                            jstop = X[neighbor + 1]
                            node = A[j]
                            while node != 0:
                                if node < 0:
                                    j = X[-node]
                                    jstop = X[-node + 1]
                                else:
                                    if marker[node] < tag:
                                        marker[node] = tag
                                        deg += supersize[node]
                                    j += 1
                                if j >= jstop:
                                    break
                                node = A[j]
                deg, mindeg = _update_external_degree(st, deg, mindeg, enode)
            enode = elimnext[enode]
        tag = mtag
        elimnode = elimnext[elimnode]
    return mindeg, tag


def _mmdnumber(neqns, invp, mergeparent):
    """SpkMMD.jl:588-633; invp is completed in place, perm returned
    (1-based list with index 0 unused)."""
    mergelastnum = [0] * (neqns + 1)
    for v in range(1, neqns + 1):
        if mergeparent[v] == 0:
            mergelastnum[v] = invp[v]
    for node in range(1, neqns + 1):
        parent = mergeparent[node]
        if parent > 0:
            root = 0
            while parent > 0:
                root = parent
                parent = mergeparent[parent]
            assert root > 0
            num = mergelastnum[root] + 1
            invp[node] = num
            mergelastnum[root] = num
            node2 = node
            while node2 != root:
                parent = mergeparent[node2]
                mergeparent[node2] = root
                node2 = parent
    perm = [0] * (neqns + 1)
    for v in range(1, neqns + 1):
        perm[invp[v]] = v
    return perm


def findetree(n, xadj, adj, rperm, rinvp):
    """_findetree! (SpkETree.jl:62-88): parent list (0 = root), 1-based
    values, returned 0-indexed by vertex - 1."""
    X = [0] + list(xadj)
    Aj = [0] + list(adj)
    rp = [0] + list(rperm)
    ri = [0] + list(rinvp)
    parent = [0] * (n + 1)
    ancestor = [0] * (n + 1)
    for i in range(1, n + 1):
        parent[i] = 0
        ancestor[i] = 0
        vertex = rp[i]
        for j in range(X[vertex], X[vertex + 1]):
            nbr = ri[Aj[j]]
            if nbr < i:
                while ancestor[nbr] != 0 and ancestor[nbr] != i:
                    nxt = ancestor[nbr]
                    ancestor[nbr] = i
                    nbr = nxt
                if ancestor[nbr] == 0:
                    parent[nbr] = i
                    ancestor[nbr] = i
    return parent[1:]


def binarytree(n, parent):
    """_binarytree! (SpkETree.jl:153-174) -> (fson, brother), 1-based lists
    with index 0 unused."""
    P = [0] + list(parent)
    fson = [0] * (n + 1)
    brother = [0] * (n + 1)
    lroot = n
    for vertex in range(n - 1, 0, -1):
        vpar = P[vertex]
        if vpar <= 0 or vpar == vertex:
            brother[lroot] = vertex
            lroot = vertex
        else:
            brother[vertex] = fson[vpar]
            fson[vpar] = vertex
    brother[lroot] = 0
    return fson, brother


def weightedbinarytree(n, parent, weight, trace=None):
    """_weightedbinarytree! (SpkETree.jl:239-274): a child whose weight is
    >= the current last son's goes to the front of its parent's list,
    otherwise it becomes the new last son (:259-268)."""
    P = [0] + list(parent)
    Wt = [0] + list(weight)
    lson = [0] * (n + 1)
    fson = [0] * (n + 1)
    brother = [0] * (n + 1)
    lroot = n
    for vertex in range(n - 1, 0, -1):
        vpar = P[vertex]
        if vpar <= 0 or vpar == vertex:
            brother[lroot] = vertex
            lroot = vertex
        else:
            vlastson = lson[vpar]
            if vlastson != 0:
                if Wt[vertex] >= Wt[vlastson]:
                    brother[vertex] = fson[vpar]
                    fson[vpar] = vertex
                    if trace is not None:
                        trace.append(("son", 261, vertex, vpar, "first"))
                else:
                    brother[vlastson] = vertex
                    lson[vpar] = vertex
                    if trace is not None:
                        trace.append(("son", 264, vertex, vpar, "last"))
            else:
                fson[vpar] = vertex
                lson[vpar] = vertex
                if trace is not None:
                    trace.append(("son", 267, vertex, vpar, "first"))
    brother[lroot] = 0
    return fson, brother


def postordertree(root, fson, brother, parent):
    """_postordertree! (SpkETree.jl:190-220) -> (invpos, new parent), 1-based
    lists with index 0 unused (parent given 1-based with index 0 unused)."""
    n = len(fson) - 1
    brother = list(brother)
    invpos = [0] * (n + 1)
    stack = [0] * (n + 1)
    num = top = 0
    vertex = root
    while vertex > 0:
        while vertex > 0:
            top += 1
            stack[top] = vertex
            vertex = fson[vertex]
        while vertex <= 0 and top > 0:
            vertex = stack[top]
            top -= 1
            num += 1
            invpos[vertex] = num
            vertex = brother[vertex]
    for vertex in range(1, num + 1):
        nuvertex = invpos[vertex]
        vpar = parent[vertex]
        if vpar > 0:
            vpar = invpos[vpar]
        brother[nuvertex] = vpar
    return invpos, brother


def getpostorder(n, parent, rperm, rinvp, weight=None, trace=None):
    """_getpostorder! (SpkETree.jl:106-139), with or without weights:
    returns (parent, rperm, rinvp, weight) after the postorder, as 1-based
    value lists (weight permuted with the vertices, or None)."""
    if weight is None:
        fson, brother = binarytree(n, parent)
    else:
        fson, brother = weightedbinarytree(n, parent, weight, trace)
    invpos, newparent = postordertree(n, fson, brother, [0] + list(parent))
    neww = None
    if weight is not None:
        stack = [0] * (n + 1)
        for v in range(1, n + 1):
            stack[invpos[v]] = weight[v - 1]
        neww = stack[1:]
    ri = [invpos[rinvp[v]] for v in range(n)]
    rp = [0] * n
    for v in range(n):
        rp[ri[v] - 1] = v + 1
    return newparent[1:], rp, ri, neww
