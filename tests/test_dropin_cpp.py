"""The C++ drop-in classes (include/gse/gse.hpp), as a C++ caller uses them.

CPU: the header compiles stand-alone with the host compiler and names every
reference class and method it replaces, the GPU ones (CudaGrid,
CudaGaussCharge, CudaConvolution) included.  GPU: tests/cpp/test_dropin (built
by __graft_entry__.build()) runs the cpu_gse.cpp testPair sequence and every
class method against libgse_hip.so and checks each result against the FP64
oracle (1e-10 relative in FP64, Erms <= 1e-5 in FP32).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HPP = os.path.join(ROOT, "include", "gse", "gse.hpp")
BIN = os.path.join(ROOT, "tests", "cpp", "test_dropin")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "gse/gse.hpp"\n'
                   "using namespace gse_amd;\n"
                   "template class CpuGaussCharge<double,double>;\n"
                   "template class CpuGaussCharge<double,float>;\n"
                   "template class CpuGaussCharge<float,float>;\n"
                   "template class CpuConvolution<float>;\n"
                   "template class CpuFFTSolver<double>;\n"
                   "template class CpuGSEk<double,double>;\n"
                   "template class CpuGSEk<float,float>;\n"
                   "template class CudaGrid<double>;\n"
                   "template class CudaGaussCharge<double,double>;\n"
                   "template class CudaGaussCharge<double,float>;\n"
                   "template class CudaGaussCharge<float,float>;\n"
                   "template class CudaConvolution<float>;\n"
                   "template class CudaConvolution<double>;\n"
                   "int main() { return 0; }\n")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)


def test_header_keeps_reference_signatures():
    txt = open(HPP).read()
    for name in ["class CpuGrid", "class CpuGaussCharge", "class CpuConvolution",
                 "class CpuGreensFunc", "class CpuGreensFuncGSE", "class CpuGreensFuncG2",
                 "class CpuFFTSolver", "class CpuGSE", "class CpuGSEk", "struct xyzq_t",
                 "int calcSupportSize(", "void calcElectricFieldOnGrid(",
                 "void interpolateForce(", "void calcDipoleSum(", "void spreadChargeToGrid(",
                 "void spreadChargeOnGrid(", "void setupKernel(", "void run(",
                 "void mul(", "void spreadCharge1(", "void spreadCharge2(",
                 "void solvePoisson()", "AT calculateEnergy()", "CpuGrid<CT>& getRhoS()",
                 "CpuGrid<CT>& getPhi()", "void setBoxSize(", "double getSigma1()",
                 "double getSigma2()", "double getLambdaSigma1()", "class CudaGrid",
                 "class CudaGaussCharge", "class CudaConvolution",
                 "void copy(const CpuGrid<T>& h_grid)", "void copy(const CudaGrid<T>& d_grid)",
                 "void clear()"]:
        assert name in txt, name


@pytest.mark.gpu
def test_cpp_dropin_program():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build()")
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120,
                       cwd=os.path.join(ROOT, "tests", "cpp"))
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "ALL PASSED" in r.stdout
