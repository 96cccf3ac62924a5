// Fixture writer compiled against the reference sources where they lie
// (never copied).  It reproduces the non-GMSH body of
// Program_Diffusion_HHO<Dim>::Execute (src/Program/Program_Diffusion_HHO.h:16-330)
// for the in-house structured meshes, with the source term of the benchmark
// problems (pi-periodic sine), and writes the assembled condensed system, the
// solver's iteration history and solution, and aggregation indices in a small
// binary format that tests/golden/make_fixtures.py packs into .npz fixtures.
//
// This program is test infrastructure: it is built only where the reference
// tree exists and is never linked into, or called by, the product.
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
#include <chrono>
#include <mutex>
#include <algorithm>
using namespace std;

#include "Utils/Utils.h"
#include "ProgramArguments.h"
#include "Discretizations/HHO/Diffusion_HHO.h"
#include "TestCases/Diffusion/DiffTestCaseFactory.h"
#include "Mesher/InHouse/Square_CartesianMesh.h"
#include "Mesher/InHouse/Square_TriangularMesh.h"
#include "Mesher/InHouse/Cube_CartesianMesh.h"
#include "Mesher/InHouse/Cube_CartesianTetrahedralMesh.h"
#include "FunctionalBasis/FunctionalBasisFactory.h"
#include "Solver/SolverFactory.h"

// ---------------------------------------------------------------- I/O helpers
static void write_csr(const string& path, const SparseMatrix& A0)
{
	SparseMatrix A = A0; // RowMajor (Types.h:19-22)
	A.makeCompressed();
	FILE* f = fopen(path.c_str(), "wb");
	int64_t hdr[3] = { (int64_t)A.rows(), (int64_t)A.cols(), (int64_t)A.nonZeros() };
	fwrite(hdr, sizeof(int64_t), 3, f);
	vector<int32_t> rp(A.rows() + 1), ci(A.nonZeros());
	for (int64_t i = 0; i <= A.rows(); i++) rp[i] = (int32_t)A.outerIndexPtr()[i];
	for (int64_t i = 0; i < A.nonZeros(); i++) ci[i] = (int32_t)A.innerIndexPtr()[i];
	fwrite(rp.data(), sizeof(int32_t), rp.size(), f);
	fwrite(ci.data(), sizeof(int32_t), ci.size(), f);
	fwrite(A.valuePtr(), sizeof(double), A.nonZeros(), f);
	fclose(f);
}

static void write_vec(const string& path, const Vector& v)
{
	FILE* f = fopen(path.c_str(), "wb");
	int64_t n = v.rows();
	fwrite(&n, sizeof(int64_t), 1, f);
	fwrite(v.data(), sizeof(double), n, f);
	fclose(f);
}

static void write_ivec(const string& path, const vector<int64_t>& v)
{
	FILE* f = fopen(path.c_str(), "wb");
	int64_t n = v.size();
	fwrite(&n, sizeof(int64_t), 1, f);
	fwrite(v.data(), sizeof(int64_t), n, f);
	fclose(f);
}

// ------------------------------------------------------------- test problems
// Benchmark source terms: -Lap u = f on the unit square/cube with
// u = prod sin(pi x_i), homogeneous Dirichlet.
static double PiSineSource2D(const DomPoint& p) { return 2 * M_PI * M_PI * sin(M_PI * p.X) * sin(M_PI * p.Y); }
static double PiSineSolution2D(const DomPoint& p) { return sin(M_PI * p.X) * sin(M_PI * p.Y); }
static double PiSineSource3D(const DomPoint& p) { return 3 * M_PI * M_PI * sin(M_PI * p.X) * sin(M_PI * p.Y) * sin(M_PI * p.Z); }
static double PiSineSolution3D(const DomPoint& p) { return sin(M_PI * p.X) * sin(M_PI * p.Y) * sin(M_PI * p.Z); }

struct Opts
{
	int dim = 2;
	string mesh = "cart";
	int n = 16;
	int k = 1;
	string source = "pisine";
	string solver = "cg";
	string prec = "mg";
	char cycle = 'V';
	int pre = 1, post = 1;
	string smoothers = "bj,bj";
	int blockSize = -1;
	double tol = 1e-8;
	int maxIter = 200;
	string out = "./out";
	bool exportMG = false;
	bool exportDat = false;
	bool aggregOnly = false;
	int aggregBlock = -1;
	bool solve = true;
	string kappaFile; // per-element isotropic kappa (int64 count + doubles, mesh element order)
	bool dumpElements = false;
	double relax = 1.0;        // Solver.RelaxationParameter (SOR smoothers)
	char nuOp = '+';           // MG.CoarseLevelChangeSmoothingOperator / Coeff
	int nuCoeff = 0;
};

// Gives every element a physical part of its own with the isotropic tensor
// kappa[i] (the reference's way of a piecewise-constant field: one named
// PhysicalGroup per part with a constant Tensor, Mesh/PhysicalGroup.h,
// TestCases/Diffusion/Tensor.h:20-23).
template <int Dim>
static void SetCellKappa(Mesh<Dim>* mesh, const string& path)
{
	FILE* f = fopen(path.c_str(), "rb");
	if (!f) { fprintf(stderr, "cannot open %s\n", path.c_str()); exit(2); }
	int64_t n = 0;
	if (fread(&n, sizeof(int64_t), 1, f) != 1 || n != (int64_t)mesh->Elements.size()) { fprintf(stderr, "kappa file size mismatch\n"); exit(2); }
	vector<double> k(n);
	if (fread(k.data(), sizeof(double), n, f) != (size_t)n) { fprintf(stderr, "short kappa file\n"); exit(2); }
	fclose(f);
	for (int64_t i = 0; i < n; i++)
	{
		PhysicalGroup<Dim>* pg = new PhysicalGroup<Dim>(1000 + (int)i, "cell" + to_string(i));
		pg->ConstantDiffTensor = new Tensor<Dim>(k[i]);
		mesh->Elements[i]->PhysicalPart = pg;
	}
}

// In-house structured meshes, as MeshFactory<Dim>::BuildMesh builds them for
// these codes (src/Mesher/MeshFactory.h:142-143,381-399).
template <int Dim> Mesh<Dim>* BuildMesh(const Opts& o);
template <> Mesh<2>* BuildMesh<2>(const Opts& o)
{
	if (o.mesh == "stri") return new Square_TriangularMesh(o.n, o.n);
	return new Square_CartesianMesh(o.n, o.n);
}
template <> Mesh<3>* BuildMesh<3>(const Opts& o)
{
	if (o.mesh == "stetra") return new Cube_CartesianTetrahedralMesh(o.n);
	return new Cube_CartesianMesh(o.n, o.n, o.n);
}

template <int Dim>
static int Run(const Opts& o)
{
	ProgramArguments args;
	args.Problem.Dimension = Dim;
	args.Problem.GeoCode = Dim == 2 ? "square" : "cube";
	args.Problem.TestCaseCode = args.Problem.GeoCode;
	args.Problem.BCCode = "d";
	args.Problem.SourceCode = o.source == "pisine" ? "sine" : o.source;
	args.Discretization.Mesher = "inhouse";
	args.Discretization.MeshCode = o.mesh;
	args.Discretization.N = o.n;
	args.Discretization.PolyDegree = o.k + 1; // -k flag: main.cpp:1061
	// Basis defaults resolved as in main.cpp:1703-1735
	bool cart = o.mesh == "cart";
	args.Discretization.OrthogonalizeElemBasesCode = cart ? 0 : 1;
	args.Discretization.ElemBasisCode = cart ? "legendre" : "monomials";
	args.Discretization.OrthogonalizeFaceBasesCode = (Dim <= 2 || cart) ? 0 : 1;
	args.Discretization.FaceBasisCode = args.Discretization.OrthogonalizeFaceBasesCode > 0 ? "monomials" : "legendre";

	args.Solver.SolverCode = o.solver;
	args.Solver.PreconditionerCode = o.prec;
	args.Solver.Tolerance = o.tol;
	args.Solver.MaxIterations = o.maxIter;
	args.Solver.PrintIterationResults = false;
	args.Solver.BlockSize = o.blockSize;
	auto& mg = args.Solver.MG;
	mg.CycleLetter = o.cycle;
	mg.PreSmoothingIterations = o.pre;
	mg.PostSmoothingIterations = o.post;
	mg.CoarseLevelChangeSmoothingOperator = o.nuOp;
	mg.CoarseLevelChangeSmoothingCoeff = o.nuCoeff;
	args.Solver.RelaxationParameter = o.relax;
	size_t comma = o.smoothers.find(',');
	mg.PreSmootherCode = o.smoothers.substr(0, comma);
	mg.PostSmootherCode = o.smoothers.substr(comma + 1);
	string p = o.prec == "default" || o.prec.empty() ? o.solver : o.prec;
	if (p == "mg")
	{
		mg.GMG_H_Prolong = GMG_H_Prolongation::CellInterp_Trace; // main.cpp:1817-1823
		if (Dim == 3 && o.mesh == "stetra")
			mg.H_CS = H_CoarsStgy::StandardCoarsening; // nested stetra hierarchy
		else
			mg.H_CS = H_CoarsStgy::StandardCoarsening;
	}
	else if (p == "aggregamg")
		mg.H_CS = H_CoarsStgy::DoublePairwiseAggregation; // main.cpp:1874-1884
	const bool uamg = p == "uamg";
	if (uamg)
	{
		// Defaults resolved by main.cpp:1847-1868 for -s uamg / -p uamg
		mg.UAMGMultigridProlong = UAMGProlongation::ChainedCoarseningProlongations;
		mg.UAMGFaceProlong = UAMGFaceProlongation::BoundaryAggregatesInteriorAverage;
		mg.UAMGCoarseningProlong = UAMGProlongation::ReconstructSmoothedTraceOrInject;
		mg.H_CS = H_CoarsStgy::MultiplePairwiseAggregation;
		mg.CoarseningFactor = 3.8;
		mg.UseGalerkinOperator = true;
	}
	args.Actions.Export.MultigridComponents = o.exportMG;
	args.OutputDirectory = o.out;
	Utils::ProgramArgs = args;

	GaussLegendre::Init();
	DiffusionTestCase<Dim>* testCase = DiffTestCaseFactory<Dim>::Create(args.Problem);
	if (o.source == "pisine")
	{
		testCase->SourceFunction = Dim == 2 ? PiSineSource2D : PiSineSource3D;
		testCase->ExactSolution = Dim == 2 ? PiSineSolution2D : PiSineSolution3D;
	}
	Mesh<Dim>* mesh = BuildMesh<Dim>(o);
	mesh->SetDiffusionField(&testCase->DiffField);
	mesh->SetBoundaryConditions(&testCase->BC);
	if (!o.kappaFile.empty())
		SetCellKappa<Dim>(mesh, o.kappaFile);

	int k = args.Discretization.PolyDegree - 1;
	FunctionalBasis<Dim>* reconstructionBasis = FunctionalBasisFactory<Dim>::Create(args.Discretization.ElemBasisCode, k + 1, false);
	FunctionalBasis<Dim>* cellBasis = FunctionalBasisFactory<Dim>::Create(args.Discretization.ElemBasisCode, k, false);
	FunctionalBasis<Dim - 1>* faceBasis = FunctionalBasisFactory<Dim - 1>::Create(args.Discretization.FaceBasisCode, k, false);
	HHOParameters<Dim>* hho = new HHOParameters<Dim>(mesh, "hho", reconstructionBasis, cellBasis, faceBasis, args.Discretization.OrthogonalizeElemBasesCode, args.Discretization.OrthogonalizeFaceBasesCode);
	// uncondensed blocks kept for U-AMG (Program_Diffusion_HHO.h:180-186)
	Diffusion_HHO<Dim>* problem = new Diffusion_HHO<Dim>(mesh, testCase, hho, true, uamg);

	auto t0 = chrono::steady_clock::now();
	args.Actions.LogAssembly = false;
	problem->Assemble(args.Actions);
	auto t1 = chrono::steady_clock::now();
	write_csr(o.out + "/A.csr", problem->A);
	write_vec(o.out + "/b.vec", problem->b);
	if (uamg)
	{
		write_csr(o.out + "/A_T_T.csr", problem->A_T_T);
		write_csr(o.out + "/A_T_F.csr", problem->A_T_ndF);
		write_csr(o.out + "/A_F_F.csr", problem->A_ndF_ndF);
	}
	if (o.exportDat)
	{
		// The reference's own Matrix Market export of the linear system
		// (Diffusion_HHO.h:682-687 -> ExportModule::ExportMatrix/ExportVector).
		ExportModule dat(o.out);
		dat.ExportMatrix(problem->A, "A");
		dat.ExportVector(problem->b, "b");
	}

	if (o.dumpElements)
	{
		// The element matrices the condensation is built from: every
		// element's local matrix Diff_HHOElement::A (cell unknowns first,
		// then its faces in element order; Diff_HHOElement.h:144,
		// FirstDOFNumber :522-525), the system number of each face (-1 for a
		// Dirichlet face, whose columns are eliminated), and the cell part of
		// the right-hand side, AssembleSourceTerm (Diffusion_HHO.h:794).
		const int nc = hho->nCellUnknowns, nfu = hho->nFaceUnknowns;
		const size_t ne = mesh->Elements.size();
		const size_t nfc = mesh->Elements[0]->Faces.size();
		const int nloc = nc + (int)nfc * nfu;
		vector<int64_t> faces(ne * nfc, -1);
		Vector loc(ne * nloc * nloc);
		for (Element<Dim>* e : mesh->Elements)
		{
			Diff_HHOElement<Dim>* he = problem->HHOElement(e);
			const size_t en = e->Number;
			for (size_t j = 0; j < nfc; j++)
				faces[en * nfc + j] = he->Faces[j]->HasDirichletBC() ? -1 : (int64_t)he->Faces[j]->Number();
			for (int r = 0; r < nloc; r++)
				for (int c = 0; c < nloc; c++)
					loc[(en * nloc + r) * nloc + c] = he->A(r, c);
		}
		write_ivec(o.out + "/elem_faces.ivec", faces);
		write_vec(o.out + "/elem_A.vec", loc);
		write_vec(o.out + "/elem_BT.vec", problem->AssembleSourceTerm(testCase->SourceFunction));
	}

	FILE* info = fopen((o.out + "/info.txt").c_str(), "w");
	fprintf(info, "rows %ld\nnnz %ld\nnf %d\nassemble_ms %.3f\n", (long)problem->A.rows(), (long)problem->A.nonZeros(), faceBasis->Size(),
		chrono::duration<double, milli>(t1 - t0).count());

	int bs = o.blockSize != -1 ? o.blockSize : faceBasis->Size();
	if (o.aggregOnly)
	{
		// Aggregation indices of the first and second pairwise passes
		// (AggregAMG.h:38-72), written per block row.
		int abs_ = o.aggregBlock > 0 ? o.aggregBlock : bs;
		AlgebraicMesh m1(abs_, 0.25);
		m1.Build(problem->A);
		bool coarsest = false;
		m1.PairWiseAggregate(coarsest);
		vector<int64_t> agg1(m1.Elements.size());
		vector<int64_t> nstrong(m1.Elements.size());
		for (size_t i = 0; i < m1.Elements.size(); i++) agg1[i] = m1.Elements[i].CoarseElement->Number;
		write_ivec(o.out + "/agg1.ivec", agg1);
		fprintf(info, "agg1_n %zu\n", m1.CoarseElements.size());
		fclose(info);
		return 0;
	}

	if (!o.solve)
	{
		fclose(info);
		return 0;
	}

	ExportModule out(o.out);
	Solver* solver = SolverFactory<Dim>::CreateSolver(args, problem, bs, out);
	IterativeSolver* it = dynamic_cast<IterativeSolver*>(solver);
	vector<double> hist;
	it->ComputeExactSolution = false;
	it->OnIterationEnd = [&hist](const IterationResult& r) { hist.push_back(r.NormalizedResidualNorm); };
	auto t2 = chrono::steady_clock::now();
	if (uamg) // Program_Diffusion_HHO.h:268-269
		it->Setup(problem->A, problem->A_T_T, problem->A_T_ndF, problem->A_ndF_ndF);
	else
		solver->Setup(problem->A);
	auto t3 = chrono::steady_clock::now();
	Vector x = it->Solve(problem->b, "0");
	auto t4 = chrono::steady_clock::now();
	write_vec(o.out + "/x.vec", x);
	if (o.exportDat) out.ExportVector(x, "solutionFaces");
	Vector h(hist.size());
	for (size_t i = 0; i < hist.size(); i++) h[i] = hist[i];
	write_vec(o.out + "/hist.vec", h);
	double l2 = -1;
	if (testCase->ExactSolution)
	{
		Vector rec = problem->ReconstructHigherOrderApproximationFromFaceCoeffs(x);
		l2 = problem->L2Error(testCase->ExactSolution, rec);
	}
	fprintf(info, "iterations %d\nsetup_ms %.3f\nsolve_ms %.3f\nl2 %.17g\n", it->IterationCount,
		chrono::duration<double, milli>(t3 - t2).count(), chrono::duration<double, milli>(t4 - t3).count(), l2);
	fclose(info);
	return 0;
}

int main(int argc, char** argv)
{
	Opts o;
	for (int i = 1; i < argc; i++)
	{
		string a = argv[i];
		auto next = [&]() { return string(argv[++i]); };
		if (a == "-dim") o.dim = stoi(next());
		else if (a == "-mesh") o.mesh = next();
		else if (a == "-n") o.n = stoi(next());
		else if (a == "-k") o.k = stoi(next());
		else if (a == "-source") o.source = next();
		else if (a == "-s") o.solver = next();
		else if (a == "-prec") o.prec = next();
		else if (a == "-cycle")
		{
			string c = next(); // e.g. V,1,1
			o.cycle = c[0];
			sscanf(c.c_str() + 2, "%d,%d", &o.pre, &o.post);
		}
		else if (a == "-smoothers") o.smoothers = next();
		else if (a == "-block-size") o.blockSize = stoi(next());
		else if (a == "-tol") o.tol = stod(next());
		else if (a == "-max-iter") o.maxIter = stoi(next());
		else if (a == "-out") o.out = next();
		else if (a == "-export-mg") o.exportMG = true;
		else if (a == "-export-dat") o.exportDat = true;
		else if (a == "-aggreg") o.aggregOnly = true;
		else if (a == "-aggreg-block") o.aggregBlock = stoi(next());
		else if (a == "-nosolve") o.solve = false;
		else if (a == "-kappa-file") o.kappaFile = next();
		else if (a == "-dump-elements") o.dumpElements = true;
		else if (a == "-relax") o.relax = stod(next());
		else if (a == "-coarse-nu") { string c = next(); o.nuOp = c[0]; o.nuCoeff = stoi(c.substr(1)); }
		else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
	}
	if (o.dim == 2) return Run<2>(o);
	return Run<3>(o);
}
