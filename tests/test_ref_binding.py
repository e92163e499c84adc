"""The reference's own functor objects bound to the MI355X solve (CPU only, no device).

Compiles tests/cpp/ref_binding_driver.cpp -- nnet_amd/reference_binding.hpp included with the
reference's headers, linked to libnnet_mi355x.so and to the reference checker build's compiled
rate functors (oracle/_ref) -- and checks that nnet::net14/86/87::computeReactionRates and
nnet::eos::IdealGasFunctor<double>(mu) / HelmholtzFunctor<double>(Z, dimension), passed as the
reference's base classes (parallel_nuclear_net.hpp:86-90) or as their own types, map to the
device parameterizations, and that functor types without a device implementation are rejected
with NNET_E_ARG.  Skipped where the reference tree or its checker build is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
LIB = os.path.join(ROOT, "sph-exa-nnet_amd", "lib", "libnnet_mi355x.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libnnet_ref.so")


def test_reference_functors_map_to_device_kinds(tmp_path):
    if not os.path.isdir(os.path.join(REFERENCE, "include", "nnet")):
        pytest.skip("reference tree not present")
    if not (os.path.exists(LIB) and os.path.exists(REF_SO)):
        pytest.skip("libnnet_mi355x.so or the reference checker build not built")
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "ref_binding_driver")
    subprocess.run([cxx, "-O1", "-std=c++20", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{REFERENCE}/include", f"-I{ROOT}/sph-exa-nnet_amd/include", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "ref_binding_driver.cpp"),
                    f"-L{os.path.dirname(REF_SO)}", "-lnnet_ref", f"-Wl,-rpath,{os.path.dirname(REF_SO)}",
                    f"-L{os.path.dirname(LIB)}", "-lnnet_mi355x", f"-Wl,-rpath,{os.path.dirname(LIB)}",
                    "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    out = subprocess.run([exe, "map"], capture_output=True, text=True, check=True).stdout.splitlines()
    print("\n".join(out))
    expect = []
    for net in (14, 86, 87):
        expect.append(f"map net{net} ideal {net} 0 10 0")  # NNET_EOS_IDEAL, mu = 10
        expect.append(f"map net{net} helmholtz {net} 1 0 {net}")  # NNET_EOS_HELMHOLTZ, z_dim = dimension
    expect += ["derived 14 86 87", "unknown rates rejected -1", "unknown eos rejected -1"]
    assert out == expect
