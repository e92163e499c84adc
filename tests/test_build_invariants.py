"""Properties of the built gfx950 code object (CPU only, no GPU).

Every LDS access of the solve kernels must stay a direct LDS instruction: the matrix fill and the
FP64 solve are lambdas forced inline with a single call site, and the single-lane rate/EOS
routines are forced inline too.  A build in which the compiler emitted the FP64 solve as a
separate function (its LDS traffic then goes through generic addresses) faulted on the device, so
no function other than the kernels may appear in the code object; the helper names stay allowed
should one of them be made a call again on purpose."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sph-exa-nnet_amd", "lib", "libnnet_mi355x.so")
LLVM = "/opt/rocm/llvm/bin"
ALLOWED_HELPERS = ("special_triple_alpha", "special_c12c12", "special_c12o16", "special_o16o16", "special_c12ag",
                   "special_o16ag", "helmholtz_cv_dpdt")


def _functions(tmp_path):
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf"):
        if not os.path.exists(os.path.join(LLVM, tool)):
            pytest.skip(f"{tool} not available")
    if not os.path.exists(LIB):
        pytest.skip("libnnet_mi355x.so not built")
    fb = str(tmp_path / "fb.bin")
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fb}", LIB, str(tmp_path / "x")],
                   check=True)
    # one offload bundle per object of the library (the host part and the solve kernels of each
    # network are separate objects); each holds its own gfx950 code object
    blob = open(fb, "rb").read()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    starts = [m.start() for m in re.finditer(re.escape(magic), blob)]
    assert starts, "no offload bundle in the library"
    funcs = {}
    for k, s in enumerate(starts):
        part, co = str(tmp_path / f"fb{k}.bin"), str(tmp_path / f"co{k}.o")
        with open(part, "wb") as f:
            f.write(blob[s:starts[k + 1] if k + 1 < len(starts) else len(blob)])
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={part}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True)
        out = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "-sW", co], check=True, capture_output=True,
                             text=True).stdout
        for line in out.splitlines():
            f = line.split()
            if len(f) >= 8 and f[3] == "FUNC":
                funcs[f[7]] = (int(f[2]), f[4])
    return funcs


def test_only_kernels_and_declared_helpers(tmp_path):
    funcs = _functions(tmp_path)
    kernels = [n for n, (_, bind) in funcs.items() if bind == "GLOBAL"]
    assert any("nnet_solve_kernel" in k for k in kernels)
    others = [n for n, (_, bind) in funcs.items() if bind != "GLOBAL"]
    unexpected = [n for n in others if not any(h in n for h in ALLOWED_HELPERS)]
    assert not unexpected, f"functions outlined from the kernels: {unexpected}"


def test_every_precision_and_network_is_built(tmp_path):
    names = set(_functions(tmp_path))
    for net in (14, 86, 87):
        for eos in (0, 1):
            for mix in (0, 1):
                pat = re.compile(rf"nnet_solve_kernelILi{net}ELi{eos}ELb{mix}E")
                assert any(pat.search(n) for n in names), (net, eos, mix)


def test_electron_capture_grids_ascend():
    """The net87 electron-capture interpolation counts the grid points below log T and below
    log(rho Ye) with one ballot per grid; that count is the index where the reference's linear
    search stops (net87.cpp:102-141) only if both grids ascend strictly."""
    import numpy as np
    text = open(os.path.join(ROOT, "sph-exa-nnet_amd", "csrc", "nnet_tables.h")).read()
    for name, n in (("NNET_ELEC_LOGT", 41), ("NNET_ELEC_LOGRHO", 51)):
        m = re.search(name + r"\[(\d+)\] = \{([^}]*)\}", text)
        assert m and int(m.group(1)) == n
        g = np.array([float(x) for x in m.group(2).split(",")])
        assert g.size == n and np.all(np.diff(g) > 0), name
        # the count of points below v equals the linear search's stopping index, also at grid points
        for v in np.concatenate([g, g + 0.05, g - 0.05, [-np.inf, np.inf, np.nan]]):
            it = 0
            while it < n and g[it] < v:
                it += 1
            assert it == int(np.sum(g < v)), (name, v)
