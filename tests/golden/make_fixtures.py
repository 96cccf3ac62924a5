"""Writes the golden fixtures of the solve path.

Each fixture is produced by ``oracle/_ref/refdrv`` (a driver of this
repository compiled against the reference sources in place, see
``oracle/Makefile``) and packed here as a compressed ``.npz``:

* ``A_*``      : the condensed operator as scalar CSR exactly as the reference
                 stores it (``Diffusion_HHO.h:633-634``, RowMajor, int32),
* ``b``, ``x`` : right-hand side and the reference solver's solution,
* ``hist``     : normalised residual after each iteration (IterationResult.h:121-139),
* ``levelL_P``/``levelL_A`` : multigrid prolongations and coarse operators
                 exported by ``Level::Setup`` (Level.h:116-125),
* ``agg1``     : first-pass pairwise aggregate of every block row
                 (AlgebraicMesh.h:54-125, PairwiseAggregation.h:9-54).
* ``ATT_*``, ``ATF_*``, ``AFF_*`` : the uncondensed blocks A_T_T, A_T_ndF,
                 A_ndF_ndF that U-AMG is set up from (Diffusion_HHO.h:52-53),
* ``uamg_factors`` : coarsening factor after every U-AMG pass, as the
                 reference prints it (UncondensedLevel.h:121-124).

* ``elem_faces``, ``elem_A``, ``elem_BT`` : every element's local matrix
                 (Diff_HHOElement::A: cell unknowns, then the faces in element
                 order), the system number of each face (-1: Dirichlet) and
                 the cell part of the right-hand side (AssembleSourceTerm),
                 the input of the batched static condensation.

* ``dat_*``    : the reference's Matrix Market export (``A.dat``, ``b.dat``,
                 ``solutionFaces.dat``; ExportModule.h:68-87) as raw bytes.

Run: ``python tests/golden/make_fixtures.py [name ...]`` (needs the reference
tree; with names, only those fixtures are regenerated).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import scipy.io as sio
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFDRV = os.path.join(REPO, "oracle", "_ref", "refdrv")
ONLY = set(sys.argv[1:])  # optional fixture names to (re)generate


def read_csr(path):
    with open(path, "rb") as f:
        rows, cols, nnz = np.frombuffer(f.read(24), dtype=np.int64)
        rp = np.frombuffer(f.read(4 * (rows + 1)), dtype=np.int32)
        ci = np.frombuffer(f.read(4 * nnz), dtype=np.int32)
        v = np.frombuffer(f.read(8 * nnz), dtype=np.float64)
    return sp.csr_matrix((v, ci, rp), shape=(int(rows), int(cols)))


def read_vec(path, dtype=np.float64):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
        return np.frombuffer(f.read(8 * n), dtype=dtype).copy()


def read_info(path):
    out = {}
    for line in open(path):
        k, v = line.split()
        out[k] = float(v)
    return out


def pack_csr(prefix, M, d):
    M = M.tocsr()
    M.sort_indices()
    d[prefix + "_rowptr"] = M.indptr.astype(np.int32)
    d[prefix + "_col"] = M.indices.astype(np.int32)
    d[prefix + "_val"] = M.data.astype(np.float64)
    d[prefix + "_shape"] = np.array(M.shape, dtype=np.int64)


def cell_kappa(dim, n, sigma, seed):
    """Per-cell lognormal kappa as this package's HHO module draws it for a
    "cart" / "hex" / "quad" mesh (the same cells in the same order as the
    reference's Cube_CartesianMesh / Square_CartesianMesh)."""
    sys.path.insert(0, REPO)
    from fhhos4_amd import fhgpu as G
    P = G.HhoProblem(dim=dim, mesh="cart", n=n, k=0, kappa="lognormal", sigma=sigma, seed=seed, defer_assembly=True)
    return P.cell_kappa()


def run(name, args, export_mg=False, keep_A=True, aggreg=None, export_dat=False, uamg=False, kappa=None,
        elements=False):
    if ONLY and name not in ONLY:
        return
    with tempfile.TemporaryDirectory() as tmp:
        cmd = [REFDRV] + args + ["-out", tmp]
        if kappa is not None:
            kf = os.path.join(tmp, "kappa.bin")
            with open(kf, "wb") as fh:
                fh.write(np.array([kappa.size], dtype=np.int64).tobytes())
                fh.write(np.ascontiguousarray(kappa, dtype=np.float64).tobytes())
            cmd += ["-kappa-file", kf]
        if export_mg:
            cmd.append("-export-mg")
        if export_dat:
            cmd.append("-export-dat")
        if elements:
            cmd.append("-dump-elements")
        out = subprocess.run(cmd, check=True, stdout=subprocess.PIPE).stdout.decode(errors="replace")
        info = read_info(os.path.join(tmp, "info.txt"))
        d = {"cmd": np.array(" ".join(["refdrv"] + args))}
        if kappa is not None:
            d["kappa"] = np.asarray(kappa, dtype=np.float64)
        if keep_A:
            pack_csr("A", read_csr(os.path.join(tmp, "A.csr")), d)
            d["b"] = read_vec(os.path.join(tmp, "b.vec"))
        if os.path.exists(os.path.join(tmp, "x.vec")):
            d["x"] = read_vec(os.path.join(tmp, "x.vec"))
            d["hist"] = read_vec(os.path.join(tmp, "hist.vec"))
        for k, v in info.items():
            d["info_" + k] = np.array(v)
        if elements:
            d["elem_faces"] = read_vec(os.path.join(tmp, "elem_faces.ivec"), dtype=np.int64).astype(np.int32)
            d["elem_A"] = read_vec(os.path.join(tmp, "elem_A.vec"))
            d["elem_BT"] = read_vec(os.path.join(tmp, "elem_BT.vec"))
        if uamg:
            for blk, f in (("ATT", "A_T_T"), ("ATF", "A_T_F"), ("AFF", "A_F_F")):
                pack_csr(blk, read_csr(os.path.join(tmp, f + ".csr")), d)
            d["uamg_factors"] = np.array([float(l.split("=")[1]) for l in out.splitlines() if "Coarsening factor" in l])
        if export_dat:
            # the reference's Matrix Market files, byte for byte
            for f in ("A", "b", "solutionFaces"):
                with open(os.path.join(tmp, f + ".dat"), "rb") as fh:
                    d["dat_" + f] = np.frombuffer(fh.read(), dtype=np.uint8)
        if export_mg:
            lvl = 0
            while os.path.exists(os.path.join(tmp, "level%d_A.dat" % lvl)):
                if lvl > 0:
                    pack_csr("level%d_A" % lvl, sio.mmread(os.path.join(tmp, "level%d_A.dat" % lvl)), d)
                pf = os.path.join(tmp, "level%d_P.dat" % lvl)
                if os.path.exists(pf):
                    pack_csr("level%d_P" % lvl, sio.mmread(pf), d)
                lvl += 1
            d["nlevels"] = np.array(lvl)
        if aggreg is not None:
            cmd = [REFDRV] + aggreg + ["-out", tmp, "-aggreg"]
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
            d["agg1"] = read_vec(os.path.join(tmp, "agg1.ivec"), dtype=np.int64).astype(np.int32)
            d["agg_cmd"] = np.array(" ".join(["refdrv"] + aggreg + ["-aggreg"]))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(name, {k: (v.shape if hasattr(v, "shape") else v) for k, v in d.items() if not k.endswith(("_col", "_val", "_rowptr"))})


def main():
    if not os.path.exists(REFDRV):
        sys.exit("build oracle/_ref/refdrv first: make -C oracle ref")
    c1 = ["-dim", "2", "-mesh", "cart", "-k", "1"]
    # Config-1 problem: 64x64 Cartesian, k=1, CG + GMG V(1,1), block Jacobi
    run("c1_n64_cg_mg_bj", c1 + ["-n", "64", "-s", "cg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bj,bj"], export_mg=True)
    run("c1_n64_cg_mg_bj23", c1 + ["-n", "64", "-s", "cg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bj23,bj23"], keep_A=False)
    # Small 2-D systems for detailed smoother / Krylov checks
    c1s = ["-dim", "2", "-mesh", "cart", "-k", "1", "-source", "exp"]
    run("c1_n32_cg_mg_bj23", c1s + ["-n", "32", "-s", "cg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bj23,bj23"], export_mg=True)
    # undamped block Jacobi stagnates here: the first iterations still pin the cycle arithmetic
    run("c1_n32_cg_mg_bj", c1s + ["-n", "32", "-s", "cg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bj,bj", "-max-iter", "15"], keep_A=False)
    run("c1_n32_cg_mg_bgs", c1s + ["-n", "32", "-s", "cg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bgs,rbgs"], keep_A=False)
    run("c1_n32_cg_bj", c1s + ["-n", "32", "-s", "cg", "-prec", "bj"], keep_A=False)
    run("c1_n32_cg_none", c1s + ["-n", "32", "-s", "cg", "-prec", ""], keep_A=False)
    run("c1_n32_fcg_mg_bj23", c1s + ["-n", "32", "-s", "fcg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bj23,bj23"], keep_A=False)
    run("c1_n32_cg_mg_w_bj23", c1s + ["-n", "32", "-s", "cg", "-prec", "mg", "-cycle", "W,1,1", "-smoothers", "bj23,bj23"], keep_A=False)
    run("c1_n32_fcg_mg_k_bgs", c1s + ["-n", "32", "-s", "fcg", "-prec", "mg", "-cycle", "K,1,1", "-smoothers", "bgs,rbgs"], keep_A=False)
    # block SOR (BlockSOR.h), alternating Gauss-Seidel, per-level smoothing change (Multigrid.h:224-242)
    run("c1_n32_cg_mg_bsor", c1s + ["-n", "32", "-s", "cg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bsor,rbsor",
                                    "-relax", "1.2"], keep_A=False)
    run("c1_n32_cg_mg_abgs", c1s + ["-n", "32", "-s", "cg", "-prec", "mg", "-cycle", "V,2,2", "-smoothers", "abgs,rabgs"],
        keep_A=False)
    # Aggregation AMG on a scalar (k=0) tetrahedral system: FCG + K-cycle
    tet0 = ["-dim", "3", "-mesh", "stetra", "-k", "0", "-n", "8", "-source", "sine"]
    run("tet8_k0_fcg_amg", tet0 + ["-s", "fcg", "-prec", "aggregamg", "-cycle", "K,1,1", "-smoothers", "bgs,rbgs"],
        export_mg=True, aggreg=tet0 + ["-nosolve"])
    run("tet8_k0_cg_amg_v", tet0 + ["-s", "cg", "-prec", "aggregamg", "-cycle", "V,1,1", "-smoothers", "bgs,rbgs", "-max-iter", "25"], keep_A=False)
    run("tet8_k0_fcg_amg_nu", tet0 + ["-s", "fcg", "-prec", "aggregamg", "-cycle", "K,1,1", "-smoothers", "bgs,rbgs",
                                      "-coarse-nu", "+1"], keep_A=False)
    # Block (n_f = 6) tetrahedral system: SpMV / block smoothers / block aggregation
    tet2 = ["-dim", "3", "-mesh", "stetra", "-k", "2", "-n", "4", "-source", "sine"]
    run("tet4_k2", tet2 + ["-s", "cg", "-prec", "bj"], aggreg=tet2 + ["-nosolve"])
    # geometric hierarchy on tetrahedra (CellInterp_Trace; the reference's orthogonalised monomial face basis)
    run("tet4_k2_cg_mg_bj23", tet2 + ["-s", "cg", "-prec", "mg", "-cycle", "V,1,1", "-smoothers", "bj23,bj23"],
        export_mg=True)
    # finer tetrahedral k=2 system: L2 error and iteration count only
    run("tet8_k2_cg_bj", ["-dim", "3", "-mesh", "stetra", "-k", "2", "-n", "8", "-source", "sine", "-s", "cg",
                          "-prec", "bj"], keep_A=False)
    # 3-D Cartesian k=2 with GMG V(1,1) bj23
    run("cart8_k2_cg_mg_bj23", ["-dim", "3", "-mesh", "cart", "-k", "2", "-n", "8", "-s", "cg", "-prec", "mg",
                                "-cycle", "V,1,1", "-smoothers", "bj23,bj23"], export_mg=True)
    # 2-D triangles (Square_TriangularMesh), k = 2: the learned-smoother family
    run("stri8_k2_cg_bj", ["-dim", "2", "-mesh", "stri", "-k", "2", "-n", "8", "-source", "pisine", "-s", "cg",
                           "-prec", "bj"])
    # Uncondensed AMG (FCG + K-cycle U-AMG, the reference's default options):
    # scalar and block hybrid systems, one with three levels
    uk = ["-s", "fcg", "-prec", "uamg", "-cycle", "K,1,1", "-smoothers", "bj23,bj23"]
    run("uamg_tet8_k0", ["-dim", "3", "-mesh", "stetra", "-k", "0", "-n", "8", "-source", "sine"] + uk,
        export_mg=True, uamg=True)
    run("uamg_tet12_k0", ["-dim", "3", "-mesh", "stetra", "-k", "0", "-n", "12", "-source", "sine"] + uk,
        export_mg=True, uamg=True)
    run("uamg_tet4_k1", ["-dim", "3", "-mesh", "stetra", "-k", "1", "-n", "4", "-source", "sine"] + uk,
        export_mg=True, uamg=True)
    run("uamg_tet4_k2", ["-dim", "3", "-mesh", "stetra", "-k", "2", "-n", "4", "-source", "sine"] + uk,
        export_mg=True, uamg=True)
    run("uamg_stri16_k2", ["-dim", "2", "-mesh", "stri", "-k", "2", "-n", "16", "-source", "pisine"] + uk,
        export_mg=True, uamg=True)
    # Cartesian cells with a per-cell lognormal kappa (one physical part per
    # cell): the unperturbed members of the hexahedral (k = 3, sigma 3) and
    # quadrilateral (k = 2, sigma 2) benchmark families
    run("hex4_k3_lognormal", ["-dim", "3", "-mesh", "cart", "-k", "3", "-n", "4", "-source", "pisine", "-s", "cg",
                              "-prec", "bj"], kappa=cell_kappa(3, 4, 3.0, 7))
    run("quad8_k2_lognormal", ["-dim", "2", "-mesh", "cart", "-k", "2", "-n", "8", "-source", "pisine", "-s", "cg",
                               "-prec", "bj"], kappa=cell_kappa(2, 8, 2.0, 5))
    # Element matrices of small systems: the input of the batched static
    # condensation and cell recovery (hexahedra with a lognormal kappa,
    # tetrahedra k = 2, quadrilaterals k = 2)
    run("elem_cart4_k1", ["-dim", "3", "-mesh", "cart", "-k", "1", "-n", "4", "-source", "pisine", "-nosolve"],
        kappa=cell_kappa(3, 4, 3.0, 7), elements=True)
    run("elem_stetra2_k2", ["-dim", "3", "-mesh", "stetra", "-k", "2", "-n", "2", "-source", "sine", "-nosolve"],
        elements=True)
    run("elem_quad4_k2", ["-dim", "2", "-mesh", "cart", "-k", "2", "-n", "4", "-source", "pisine", "-nosolve"],
        elements=True)
    # the headline cells: hexahedra k = 3 (60 face unknowns per cell: the MFMA Schur product)
    run("elem_cart3_k3", ["-dim", "3", "-mesh", "cart", "-k", "3", "-n", "3", "-source", "pisine", "-nosolve"],
        kappa=cell_kappa(3, 3, 3.0, 9), elements=True)
    # Exported linear systems (Matrix Market text) of tiny 2-D and 3-D problems
    run("mm_c1_n4", c1s + ["-n", "4", "-s", "cg", "-prec", "mg", "-smoothers", "bj,bj"], export_dat=True)
    run("mm_tet2_k1", ["-dim", "3", "-mesh", "stetra", "-k", "1", "-n", "2", "-source", "sine", "-s", "cg",
                       "-prec", "bj"], export_dat=True)


if __name__ == "__main__":
    main()
