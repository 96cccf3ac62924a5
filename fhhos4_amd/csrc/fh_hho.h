// HHO discretisation of -div(K grad u) = f with homogeneous Dirichlet
// conditions on structured affine meshes, statically condensed onto the face
// unknowns: the producer of the face system A x_F = b that the solvers take.
//
// It follows the reference's HHO operators (src/Discretizations/HHO/
// Diff_HHOElement.h:104-509, Diffusion_HHO.h:252-704): reconstruction in
// P^{k+1}(T) with the mean-value condition, consistency P^T S P, the "hho"
// stabilisation with weight (K n.n)/h_F, static condensation
// A = A_FF - A_TF^T A_TT^-1 A_TF, b = -A_TF^T A_TT^-1 B_T, and the
// relative L2 error of the reconstructed solution (Diffusion_HHO.h:120-149).
// Bases are orthonormal in L2 of each cell / face (built once on the
// reference shape; the reference uses Legendre or orthogonalised monomials,
// which span the same spaces, so the discrete solution is the same).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "fh_common.h"
#include "fh_uamg.h"

namespace fh {

struct HhoArgs {
    int dim = 3;
    std::string mesh = "stetra";   // "stetra" (cube split into 6 tets) | "cart" (squares / cubes) | "stri" (2-D triangles)
                                   // | "hex" / "quad" (trilinear / bilinear cells of the grid, perturbable)
    int n = 8;                     // cells per direction of the background grid
    int k = 2;                     // face degree
    double perturb = 0.0;          // interior vertex perturbation, fraction of h (uniform per coordinate)
    uint64_t seed = 1;
    std::string source = "pisine"; // "pisine": f = d pi^2 prod sin(pi x) | "sine": 4 pi periodic |
                                   // "modes": sum of source_modes random modes a prod sin(p_d pi x_d)
    uint64_t source_seed = 1;
    int source_modes = 8;
    std::string kappa = "one";     // "one" | "lognormal" (per cell) | "subdomain" (per block of cells)
    double sigma = 0.0;            // lognormal standard deviation
    int subdomain = 8;             // cells per subdomain edge ("subdomain"), counted on the reference grid
    double aniso = 1e-2;           // z-scaling of the anisotropic subdomains ("subdomain")
    bool host_assembly = true;     // false: only mesh, tables and pattern on the host (AssembleDevice)
    int ref_n = 0;                 // finest grid of the hierarchy (0: n).  Vertex perturbations and
                                   // subdomains are defined on it, so a mesh with n = ref_n / 2^l
                                   // is the geometric coarsening of the finest one.
};

struct HhoTimes {
    double mesh_s = 0, local_s = 0, assemble_s = 0;
};

class HhoProblem {
public:
    explicit HhoProblem(const HhoArgs& a);
    ~HhoProblem();
    // condensed system: nb block rows (interior faces) of nf x nf blocks
    int nb = 0, nf = 0;
    std::vector<int> rowptr, colidx;
    std::vector<double> vals, b;
    int ncells() const;
    int nfaces() const;
    // relative L2 error of the k+1 reconstruction of the solution x (face
    // unknowns, length nb*nf) against the exact solution of the source
    double L2Error(const double* x) const;
    // The same on the GPU: x_dev is the face solution in device memory; the
    // cell recovery, k+1 reconstruction and quadrature run one wavefront per
    // cell (Diffusion_HHO.h:120-149, 855-913, 960-966).
    double L2ErrorDevice(Ctx& c, const double* x_dev) const;
    // cell diffusion coefficient (isotropic part) of every cell
    std::vector<double> CellKappa() const;
    // face adjacency: cells of every interior face (2 per face)
    void FaceCells(std::vector<int>& out) const;
    // Prolongation from the problem on the once-coarsened mesh (n/2) to this
    // one, as nf x nf blocks (fine faces x coarse faces): the coarse cell's
    // k+1 reconstruction from its face values, L2-projected on the fine
    // faces, weighted by normal diffusivity across coarse faces
    // (CellInterp_Trace, LevelForHHO.h:192-213, 743-763, 899-954).
    void Prolongation(const HhoProblem& coarse, std::vector<int>& prp, std::vector<int>& pci,
                      std::vector<double>& pv) const;
    // Uncondensed hybrid blocks, laid out as the reference keeps them for
    // U-AMG (Diffusion_HHO::A_T_T, A_T_ndF, A_ndF_ndF; Diffusion_HHO.h:52-53,
    // 410-470, 564-624): cells in mesh order with cell_unknowns() DOFs each,
    // faces in block-row order.  Entries are in this module's local bases
    // (orthonormal on the reference shape), so for k >= 1 they equal the
    // reference's up to a change of the cell and face bases; their Schur
    // complement is the condensed operator.
    void Uncondensed(Csr& ATT, Csr& ATF, Csr& AFF) const;
    int cell_unknowns() const;
    HhoTimes times;
    // mesh perturbation: rejected draws, vertices left on their grid
    // position, smallest cell quality (Jacobian determinant / h^dim)
    int64_t redraws() const;
    int64_t kept_vertices() const;
    double min_quality() const;
    bool host_assembled = false;       // vals and b hold the host assembly
    mutable double device_assemble_s = 0;
    // Local operators, static condensation and assembly on the GPU (one
    // workgroup per cell): the block operator on the device (the host
    // pattern rowptr / colidx) and the right-hand side into b_dev (nb*nf).
    std::unique_ptr<Bcsr> AssembleDevice(Ctx& c, double* b_dev, bool keep_host) const;
    struct Impl;
    Impl* impl = nullptr;
};

}  // namespace fh
