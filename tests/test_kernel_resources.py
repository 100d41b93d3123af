"""Every kernel of the built library runs without scratch memory: a kernel
whose registers spill to scratch runs many times slower (the backward block
GEMV once did when it read eight right-hand sides per pass). The figures are
the compiler's own, read from the gfx950 code object's metadata
(NT_AMDGPU_METADATA: .private_segment_fixed_size and .vgpr_count per kernel)
of the library built from this repository's sources."""
import os
import re
import struct
import subprocess

import pytest

from sparspak_hip import _lib

READELF = "/opt/rocm/llvm/bin/llvm-readelf"


def _kernel_metadata(tmp_path):
    data = open(_lib.LIB_PATH, "rb").read()
    base = data.find(b"__CLANG_OFFLOAD_BUNDLE__")
    assert base >= 0, "no offload bundle in the library"
    p = base + 24
    (n,) = struct.unpack_from("<Q", data, p)
    p += 8
    for _ in range(n):
        off, size, tl = struct.unpack_from("<QQQ", data, p)
        p += 24
        triple = data[p:p + tl].decode()
        p += tl
        if triple.endswith("gfx950"):
            co = tmp_path / "gfx950.co"
            co.write_bytes(data[base + off:base + off + size])
            return subprocess.run([READELF, "--notes", str(co)], capture_output=True,
                                  text=True, check=True).stdout
    raise AssertionError("no gfx950 code object in the library")


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
@pytest.mark.skipif(not os.path.exists(READELF), reason="llvm-readelf not available")
def test_no_kernel_spills_to_scratch(tmp_path):
    meta = _kernel_metadata(tmp_path)
    kernels = re.split(r"\n  - \.", meta.split("amdhsa.kernels:", 1)[1])
    found = {}
    for k in kernels:
        name = re.search(r"\.name:\s+(\S+)", k)
        scratch = re.search(r"\.private_segment_fixed_size:\s+(\d+)", k)
        vgpr = re.search(r"\.vgpr_count:\s+(\d+)", k)
        if name and scratch:
            found[name.group(1)] = (int(scratch.group(1)), int(vgpr.group(1)) if vgpr else -1)
    assert len(found) > 40, f"only {len(found)} kernels found in the metadata"
    spills = {k: v for k, v in found.items() if v[0] != 0}
    assert not spills, f"kernels using scratch (bytes, vgprs): {spills}"
