/*
 * z4c_oracle.h -- CPU restatement of the GR-Athena++ Z4c evolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never links or calls it.
 *
 * It restates, in plain C99 with IEEE double arithmetic and no FMA
 * contraction (built with -ffp-contract=off), the per-point arithmetic of:
 *   src/utils/finite_differencing.tpp          (Dx, Dxx, Ds, Lx, Diss; the
 *                                               DBG_SYMMETRIZE_FD branches,
 *                                               which src/defs.hpp.in:18 enables)
 *   src/utils/finite_differencing_stencils.tpp (coefficient tables)
 *   src/z4c/calculate_z4c_rhs.cpp:29-916      (Z4c::Z4cRHS)
 *   src/z4c/adm_z4c.cpp:972-1120              (Z4c::PrepareZ4cDerivatives)
 *   src/z4c/z4c.cpp:677-991                   (Z4c::AlgConstr, both forms)
 *   src/z4c/adm_z4c.cpp:46-365                (ADMToZ4c, Z4cToADM)
 *   src/z4c/adm_z4c.cpp:382-930               (ADMConstraints)
 *   src/z4c/boundary_conditions.cpp:16-174    (Z4cBoundaryRHS / Sommerfeld)
 *   src/task_list/task_list_integrators_butcher.cpp:542-680 (SumBT_ak/bk)
 *   src/z4c/add_z4c_rhs.cpp:24-34, src/mesh/weighted_ave.cpp:157-167
 *   src/z4c/new_blockdt_z4c.cpp:24-70
 *
 * Layout: every field array is AthenaArray order u(n,k,j,i) =
 * data[i + nn1*(j + nn2*(k + nn3*n))] (src/athena_arrays.hpp:192-193).
 * Z4c variable order follows src/z4c/z4c.hpp:54-88.
 */
#ifndef Z4C_ORACLE_H_
#define Z4C_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Z4c::opt (src/z4c/z4c.hpp:374-430, defaults src/z4c/z4c.cpp:318-472). */
typedef struct {
  double chi_psi_power;
  double chi_div_floor;
  double diss;
  double eps_floor;
  double damp_kappa1;
  double damp_kappa2;
  double lapse_oplog;
  double lapse_harmonicf;
  double lapse_harmonic;
  double lapse_advect;
  double lapse_K;
  double shift_Gamma;
  double shift_alpha2Gamma;
  double shift_H;
  double shift_advect;
  double shift_eta;
  double r_max_con;
  int cowling;
  int tf_force_regularization;
  int force_regularization;
  int pad_;
} oracle_params;

/* Vertex-centred MeshBlock geometry (src/mesh/meshblock.cpp:852-915). */
typedef struct {
  int nn1, nn2, nn3;             /* padded vertex counts                 */
  int ng;                        /* ghost width, must be 4               */
  int il, iu, jl, ju, kl, ku;    /* interior (owned + shared) vertices   */
  double dx[3];                  /* uniform spacing per axis             */
  double x0[3];                  /* coordinate of padded index 0         */
} oracle_geom;

enum {
  OR_CHI = 0, OR_GXX = 1, OR_KHAT = 7, OR_AXX = 8, OR_GAMX = 14,
  OR_THETA = 17, OR_ALPHA = 18, OR_BETAX = 19, OR_NZ4C = 22
};
enum {
  OR_ADM_ALPHA = 0, OR_ADM_BETAX = 1, OR_ADM_GXX = 4, OR_ADM_KXX = 10,
  OR_ADM_PSI4 = 16, OR_NADM = 17
};
enum {
  OR_CON_C = 0, OR_CON_H = 1, OR_CON_M = 2, OR_CON_Z = 3, OR_CON_MX = 4,
  OR_NCON = 7
};
enum { OR_MAT_RHO = 0, OR_MAT_SX = 1, OR_MAT_SXX = 4, OR_NMAT = 10 };
/* Prepared first derivatives: d1[b*11 + f], f = alpha, chi, beta^x..z,
   g~xx,g~xy,g~xz,g~yy,g~yz,g~zz; plus ddchi[6] (sym order). */
enum { OR_ND1 = 33, OR_NDDCHI = 6, OR_NAUX = 30, OR_NAUXEXT = 7 };

/* Single-stencil application at flat index p along axis dir (0,1,2).
   kind: 0 Dx, 1 Dxx, 2 Ds, 3 Lx-left, 4 Lx-right, 5 Diss (diss=1).
   Lx-* return the raw biased derivative (vx = 1). */
double oracle_fd_point(int kind, const oracle_geom* g, const double* u,
                       long p, int dir);
/* Lx(dir, vx, u) at flat index p. */
double oracle_lx_point(const oracle_geom* g, const double* u, long p, int dir,
                       double vx);

void oracle_prep_derivs(const oracle_params* o, const oracle_geom* g,
                        const double* u, double* d1, double* ddchi,
                        double* aux /* may be NULL */);

void oracle_rhs(const oracle_params* o, const oracle_geom* g, const double* u,
                const double* mat /* may be NULL (vacuum) */,
                const double* d1, const double* ddchi, double* rhs);

/* face flags: 6 ints (ix1, ox1, ix2, ox2, ix3, ox3), nonzero = Sommerfeld. */
void oracle_boundary_rhs(const oracle_params* o, const oracle_geom* g,
                         const double* u, const int* sommerfeld_faces,
                         double* rhs);

void oracle_algconstr(const oracle_params* o, const oracle_geom* g, double* u);
void oracle_algconstr_range(const oracle_params* o, const oracle_geom* g,
                            double* u, int il, int iu, int jl, int ju, int kl,
                            int ku, int skip_physical);

void oracle_adm_to_z4c(const oracle_params* o, const oracle_geom* g,
                       const double* adm, double* u,
                       double* aux_ext /* may be NULL */);
void oracle_z4c_to_adm(const oracle_params* o, const oracle_geom* g,
                       const double* u, double* adm,
                       double* aux_ext /* may be NULL */);

void oracle_adm_constraints(const oracle_params* o, const oracle_geom* g,
                            const double* u, const double* adm,
                            const double* mat /* may be NULL */,
                            const double* d1, const double* ddchi, double* con);

/* u_out = u0 + sum_j (w_j*dt) k_j over the whole padded block, skipping
   zero weights, in j order (SumBT_ak / SumBT_bk). */
void oracle_bt_sum(const oracle_geom* g, int nvar, double* u_out,
                   const double* u0, int nk, const double* const* k,
                   const double* w, double dt);
/* WeightedAve (WeightedAveVC: owned vertices is..ive) and AddZ4cRHS
   (owned vertices). */
void oracle_weighted_ave(const oracle_geom* g, int nvar, double* u_out,
                         const double* u_in1, const double* u_in2,
                         const double* wght);
void oracle_add_rhs(const oracle_geom* g, int nvar, const double* rhs,
                    double wght, double* u_out);
double oracle_new_dt(const oracle_geom* g, double cfl);

/* Z4c columns of the history output (src/outputs/history.cpp:178-205) of
   one block: out[8] += vol*{H^2,|M|,Mx^2,My^2,Mz^2,|Z|,Theta^2,|C|} over the
   cell-index loop (vertices il..iu-1 etc.), cell centres within r_max_con. */
void oracle_history(const oracle_params* o, const oracle_geom* g,
                    const double* u, const double* con, double* out);

/* Same-level VC exchange over nb blocks of one spacing on the domain
   [lo, hi] (periodic[d] != 0: axis d wraps): every vertex with at least
   one owner (an interior vertex of any block at the same lattice point)
   takes the owners' mean; u[b] is block b's nvar-field array. */
void oracle_exchange(int nb, const oracle_geom* g, double* const* u, int nvar,
                     const double* lo, const double* hi, const int* periodic);

/* Physical ghost fill of one block on the flagged faces (6 flags,
   ix1..ox3): 5-point extrapolation as the reference applies to VC data on
   gr_sommerfeld / extrapolate_outflow faces, with its cell-centred ranges. */
void oracle_physical_bc(const oracle_geom* g, int nvar, double* u, const int* faces);

/* Single-block periodic vertex-centred ghost fill (all three axes). */
void oracle_periodic_fill(const oracle_geom* g, int nvar, double* u);

/* Z4c::Z4cWeyl (src/z4c/calculate_weyl_scalars.cpp:44-628): out[2][N] =
   Re, Im psi4 at interior vertices (NaN elsewhere) from the ADM variables
   adm[17][N], the state and the prepared derivatives of oracle_prep_derivs. */
void oracle_weyl(const oracle_params* o, const oracle_geom* g, const double* u,
                 const double* adm, const double* d1, const double* ddchi, double* out);

/* VC refinement (src/mesh/mesh_refinement_vc.cpp; comm/refinement_ops.*):
   cg is the coarse geometry of block g (oracle.Geom.coarse): per axis
   nx/2 + 9 vertices, interior 4 .. 4 + nx/2.  Restriction is
   RestrictVertexCenteredValues over the coarse interior (RestrictInterior);
   prolongation is ProlongateVertexCenteredValues (3-D, even NGHOST) over the
   coarse index box [si..ei] x [sj..ej] x [sk..ek]; prolongation_range is
   ProlongationRange (VC) for one axis of a coarser neighbour at offset ox
   (parity: the block's logical coordinate & 1). */
void oracle_restrict_vc(const oracle_geom* g, const oracle_geom* cg, int nvar,
                        const double* fine, double* coarse);
void oracle_prolongate_vc(const oracle_geom* g, const oracle_geom* cg, int nvar,
                          const double* coarse, double* fine, int si, int ei, int sj,
                          int ej, int sk, int ek);
void oracle_prolongation_range(int parity, int ox, int cvs, int cve, int* s, int* e);

/* Restart-file records as RestartOutput::WriteOutputFile copies them
   (src/outputs/restart.cpp:141-147, 206-214): the mesh header (nbtotal,
   root_level, RegionSize, time, dt, ncycle, datasize) and one ID-list record
   (LogicalLocation, cost), the structs declared as src/athena.hpp:58-96
   declares them, padding zeroed.  Return the bytes written. */
unsigned long oracle_rst_header(unsigned char* out, int nbtotal, int root_level,
                                const double* size9, const int* nx, double time, double dt,
                                int ncycle, unsigned long long datasize);
unsigned long oracle_rst_idrec(unsigned char* out, const long long* lx, int level,
                               double cost);

#ifdef __cplusplus
}
#endif
#endif /* Z4C_ORACLE_H_ */
