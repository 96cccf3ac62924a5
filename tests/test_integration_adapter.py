"""The reference-side adapter printed in INTEGRATION.md (the GpuSolver class
and the SolverFactory branch) compiles against the reference's own headers
(src/Solver/IterativeSolver.h, src/Solver/SolverFactory.h,
src/ProgramArguments.h) and this package's C ABI (include/fhgpu.h).

Only where the reference tree exists (this is a check of the documentation's
code, not of the library)."""
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
EIGEN = os.path.join(REF, "dependencies", "eigen", "eigen-eigen-b3f3d4950030")
ROCM_INC = "/opt/rocm/include"

PROLOGUE = """
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <map>
#include <list>
#include <set>
#include <functional>
#include <iostream>
#include <fstream>
#include <mutex>
#include <algorithm>
using namespace std;
#include "Utils/Utils.h"
#include "ProgramArguments.h"
#include "Discretizations/HHO/Diffusion_HHO.h"
#include "Solver/SolverFactory.h"
#include "GpuSolver.h"
"""

FACTORY = """
// the factory branch in the context of SolverFactory<Dim>::CreateSolver
// (SolverFactory.h:126): args, problem, blockSize and solver in scope
template <int Dim>
Solver* CreateWithGpuBranch(const ProgramArguments& args, Diffusion_HHO<Dim>* problem, int blockSize)
{
	Solver* solver = nullptr;
	if (args.Solver.SolverCode.compare("cg") == 0)
		solver = nullptr;
%s
	return solver;
}
template Solver* CreateWithGpuBranch<2>(const ProgramArguments&, Diffusion_HHO<2>*, int);
template Solver* CreateWithGpuBranch<3>(const ProgramArguments&, Diffusion_HHO<3>*, int);

// the calls the reference makes through IterativeSolver
void Drive(GpuSolver& g, const SparseMatrix& A, const Vector& b)
{
	g.Setup(A);
	Vector x0 = g.Solve(b);
	Vector x1 = static_cast<IterativeSolver&>(g).Solve(b, "rand");
	Vector x2 = x1;
	g.Solve(b, x2, false, true, true);
	(void)x0;
}
"""


def _cpp_blocks():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def test_integration_notes_hold_the_adapter_and_the_factory_branch():
    blocks = _cpp_blocks()
    assert len(blocks) >= 2
    assert "class GpuSolver : public IterativeSolver" in blocks[0]
    assert "new GpuSolver(" in blocks[1]
    # every fh_ entry point the adapter calls is declared in the C ABI
    header = open(os.path.join(REPO, "include", "fhgpu.h")).read()
    for name in set(re.findall(r"\b(fh_[a-z0-9_]+)\(", blocks[0] + blocks[1])):
        assert re.search(r"\b%s\(" % name, header), name


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree absent")
@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(ROCM_INC), reason="g++ or ROCm headers absent")
def test_adapter_compiles_against_the_reference_headers(tmp_path):
    cls, branch = _cpp_blocks()[:2]
    (tmp_path / "GpuSolver.h").write_text(cls)
    (tmp_path / "tu.cpp").write_text(PROLOGUE + FACTORY % branch)
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-w", "-D__HIP_PLATFORM_AMD__", "-DSMALL_INDEX", "-DENABLE_2D",
           "-DENABLE_3D", "-I", str(tmp_path), "-I", os.path.join(REPO, "include"), "-I", ROCM_INC,
           "-I", os.path.join(REF, "src"), "-I", os.path.join(REF, "src", "Solver"), "-I", EIGEN,
           str(tmp_path / "tu.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
