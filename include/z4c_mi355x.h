/*
 * z4c_mi355x.h -- C-ABI of the MI355X (gfx950) Z4c evolution hot path.
 *
 * Drop-in for the per-MeshBlock Z4c class surface of GR-Athena++
 * (src/z4c/z4c.hpp:448-610) on vertex-centred MeshBlocks, batched: every
 * call acts on all blocks registered in the context in one launch.
 *
 * Conventions
 *  - All field pointers are DEVICE pointers (hipMalloc'd, FP64).  A "state
 *    buffer" holds every registered block back to back; block b starts at
 *    element offset blocks[b].offset and stores nvar fields in the
 *    reference's AthenaArray order u(n,k,j,i) = base[i + nn1*(j + nn2*(k +
 *    nn3*n))] (src/athena_arrays.hpp:192-193).  The same offsets index every
 *    state-shaped buffer (u, rhs, RK registers, adm, con, mat).
 *  - Z4c variable order chi, g~(6), Khat, A~(6), Gam~(3), Theta, alpha,
 *    beta(3) (src/z4c/z4c.hpp:54-88); ADM order src/z4c/z4c.hpp:99-135;
 *    constraint order C,H,M,Z,Mx,My,Mz (src/z4c/z4c.hpp:140-160).
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Calls are stream ordered and never synchronise the host unless stated.
 *  - Every function returns 0 on success or a negative Z4C_E* code; the
 *    message is available from z4c_last_error().  Nothing throws across the
 *    ABI.  A context is used by one host thread at a time (the reference's
 *    Z4c objects are not re-entrant either, src/z4c/z4c.hpp:613-698).
 */
#ifndef Z4C_MI355X_H_
#define Z4C_MI355X_H_

#ifdef __cplusplus
extern "C" {
#endif

#define Z4C_ABI_VERSION 1

enum {
  Z4C_OK = 0,
  Z4C_EINVAL = -1,   /* bad argument (null pointer, size, range)   */
  Z4C_EHIP = -2,     /* a HIP runtime call failed                   */
  Z4C_ESTATE = -3,   /* call out of order (e.g. no blocks set)      */
  Z4C_ENOMEM = -4    /* device allocation failed                    */
};

enum { Z4C_NVAR = 22, Z4C_NADM = 17, Z4C_NCON = 7, Z4C_NMAT = 10 };

/* Z4c::opt (src/z4c/z4c.hpp:374-430); defaults src/z4c/z4c.cpp:318-472. */
typedef struct {
  double chi_psi_power;      /* -4 */
  double chi_div_floor;      /* -1000 */
  double diss;               /* 0 */
  double eps_floor;          /* 1e-12 */
  double damp_kappa1;        /* 0 */
  double damp_kappa2;        /* 0 */
  double lapse_oplog;        /* 2 */
  double lapse_harmonicf;    /* 1 */
  double lapse_harmonic;     /* 0 */
  double lapse_advect;       /* 1 */
  double lapse_K;            /* 0 */
  double shift_Gamma;        /* 1 */
  double shift_alpha2Gamma;  /* 0 */
  double shift_H;            /* 0 */
  double shift_advect;       /* 1 */
  double shift_eta;          /* 2 */
  double r_max_con;          /* +inf */
  int cowling;               /* 0 */
  int tf_force_regularization; /* 1 */
  int force_regularization;  /* 0 */
  int pad_;
} z4c_params;

/* One vertex-centred MeshBlock (MB_info, src/mesh/mesh.hpp:817-829, and the
   VC ranges of src/mesh/meshblock.cpp:869-889): nn = nx + 1 + 2*4. */
typedef struct {
  long long offset;            /* element offset of the block in a state buffer */
  int nn1, nn2, nn3;           /* padded vertex counts per axis                 */
  int il, iu, jl, ju, kl, ku;  /* interior vertex ranges (inclusive)            */
  double dx[3];                /* uniform spacing                               */
  double x0[3];                /* coordinate of padded index 0 on each axis     */
  int level;                   /* refinement level (bookkeeping only)           */
  int sommerfeld_faces;        /* bit f set: face f (ix1,ox1,ix2,ox2,ix3,ox3)
                                  carries gr_sommerfeld (Z4cBoundaryRHS and the
                                  physical ghost fill)                          */
  int outflow_faces;           /* bit f set: face f carries extrapolate_outflow
                                  (physical ghost fill only)                    */
  int pad_;
} z4c_block;

typedef struct z4c_ctx z4c_ctx;

/* Last error message of the calling thread (static storage). */
const char* z4c_last_error(void);
int z4c_abi_version(void);

/* Context on HIP device `device`.  Replaces the per-block Z4c constructor's
   allocation of derivative storage (src/z4c/z4c.cpp:475-550). */
int z4c_ctx_create(z4c_ctx** ctx, int device);
int z4c_ctx_destroy(z4c_ctx* ctx);

/* Register the block list (host array).  Allocates the per-block prepared
   derivative storage and the launch tables.  Replaces the MeshBlock loop of
   TaskList::DoTaskListOneStage (src/task_list/task_list.cpp:122-261).
   Returns the number of doubles one state buffer of `nvar` fields needs in
   *elems_per_var (sum of nn1*nn2*nn3 over blocks; multiply by nvar). */
int z4c_set_blocks(z4c_ctx* ctx, const z4c_block* blocks, int nblocks,
                   long long* elems_per_var);

/* Z4c::PrepareZ4cDerivatives (src/z4c/adm_z4c.cpp:972-1120): first
   derivatives of alpha, chi, beta^i, g~_ij on the extended ranges, kept in
   the context.  `aux` (may be NULL) receives the 30 physical ADM
   derivatives (storage.aux, interior).  The right-hand side does not need
   them here (its stencil pass forms every derivative it uses) and
   z4c_adm_constraints prepares its own. */
int z4c_prepare_derivatives(z4c_ctx* ctx, const z4c_params* p, const double* u,
                            double* aux, void* stream);

/* Z4c::Z4cRHS (src/z4c/calculate_z4c_rhs.cpp:29-916) incl. KO dissipation,
   interior points only; all derivatives are formed from `u` (ghosts
   included).  `mat` (10 fields rho, S_a, S_ab, src/z4c/z4c.hpp:163-184)
   may be NULL (vacuum).  Ghosts of rhs are not written. */
int z4c_rhs(z4c_ctx* ctx, const z4c_params* p, const double* u,
            const double* mat, double* rhs, void* stream);

/* Z4c::Z4cBoundaryRHS (src/z4c/boundary_conditions.cpp:16-174): Sommerfeld
   conditions on the faces flagged in z4c_block.sommerfeld_faces. */
int z4c_boundary_rhs(z4c_ctx* ctx, const z4c_params* p, const double* u,
                     double* rhs, void* stream);

/* Z4c::AlgConstr (src/z4c/z4c.cpp:677-799), whole padded block, in place. */
int z4c_algconstr(z4c_ctx* ctx, const z4c_params* p, double* u, void* stream);

/* Z4c::AlgConstr ranged form (src/z4c/z4c.cpp:801-991): the same index
   ranges on every block; skip_physical != 0 leaves interior points alone. */
int z4c_algconstr_range(z4c_ctx* ctx, const z4c_params* p, double* u, int il,
                        int iu, int jl, int ju, int kl, int ku,
                        int skip_physical, void* stream);

/* Z4c::ADMToZ4c (src/z4c/adm_z4c.cpp:46-192) incl. the closing AlgConstr.
   `aux_ext` (7 fields: sqrt det gamma, gamma^ij; may be NULL) receives
   PrepareAuxExtended (src/z4c/adm_z4c.cpp:1139-1231). */
int z4c_adm_to_z4c(z4c_ctx* ctx, const z4c_params* p, const double* adm,
                   double* u, double* aux_ext, void* stream);

/* Z4c::Z4cToADM (src/z4c/adm_z4c.cpp:200-260), whole block. */
int z4c_z4c_to_adm(z4c_ctx* ctx, const z4c_params* p, const double* u,
                   double* adm, double* aux_ext, void* stream);

/* Z4c::ADMConstraints (src/z4c/adm_z4c.cpp:382-930): con (7 fields) zeroed
   then set on the interior.  First prepares the derivatives of `u`
   (PrepareZ4cDerivatives), as PREP_Z4C_DERIV precedes ADM_CONSTR in the
   task list (src/task_list/gr/task_list_gr_z4c.cpp:67-71). */
int z4c_adm_constraints(z4c_ctx* ctx, const z4c_params* p, const double* u,
                        const double* adm, const double* mat, double* con,
                        void* stream);

/* Butcher::SumBT_ak / SumBT_bk (src/task_list/task_list_integrators_butcher.cpp:573-680):
   u_out = u0 + sum_{j<nk, w[j]!=0} (w[j]*dt) k[j], whole padded block, j in
   order.  `k` is a HOST array of nk device pointers; nk <= 8.  u_out may
   alias u0. */
int z4c_bt_sum(z4c_ctx* ctx, int nvar, double* u_out, const double* u0, int nk,
               const double* const* k, const double* w, double dt,
               void* stream);

/* MeshBlock::WeightedAveVC (src/mesh/weighted_ave.cpp:21-167, interior
   vertices) and Z4c::AddZ4cRHS (src/z4c/add_z4c_rhs.cpp:24-34). */
int z4c_weighted_ave(z4c_ctx* ctx, int nvar, double* u_out, const double* u_in1,
                     const double* u_in2, const double* wght3, void* stream);
int z4c_add_rhs(z4c_ctx* ctx, int nvar, const double* rhs, double wght,
                double* u_out, void* stream);

/* Z4c::NewBlockTimeStep (src/z4c/new_blockdt_z4c.cpp:24-70): minimum over
   registered blocks of cfl*min(dx); host-side, no device work. */
int z4c_new_dt(z4c_ctx* ctx, double cfl, double* dt_out);

/* Single-block periodic vertex-centred ghost exchange (the periodic
   self-neighbour case of CommRegistry::SendBoundaryBuffers/SetBoundaries with
   additive VC unpack and node-multiplicity division,
   src/comm/comm_channel.cpp:1277-1330, src/comm/index_utilities.hpp:116-139).
   Applies to every registered block independently. */
int z4c_periodic_fill(z4c_ctx* ctx, int nvar, double* u, void* stream);

/* Vertex-centred ghost update between the registered blocks, as GR_Z4c's
   SEND_Z4C / RECV_Z4C / SETB_Z4C (and PROLONG across levels) do it
   (src/task_list/gr/task_list_gr_z4c.cpp:47-58).  z4c_set_exchange builds
   the update for the domain [lo, hi] (periodic[d] != 0: axis d wraps) once
   per block list, with its scratch for up to 22 fields; z4c_exchange
   applies it (stream-ordered; with at most 22 fields it allocates nothing
   and can be captured into a HIP graph).
   * All blocks on one spacing (sitting on its lattice): every vertex that
     some block owns as an interior vertex (ghosts, and shared boundary
     vertices) becomes the mean of its owners' values, summed in block
     order and k, j, i order, times 1/multiplicity (additive unpack
     src/comm/comm_channel.cpp:1277-1330, ranges
     src/comm/index_utilities.hpp:116-139, division
     src/comm/node_multiplicity.cpp:349-420).
   * Blocks on several spacings (equal cell counts nx >= 8, even, 4 ghost
     layers; spacings h0 * 2^-L on a 2:1-balanced block tree of the domain):
     the reference's multi-level update.  Neighbours as
     mesh_topology::SearchAndSetNeighbors finds them
     (src/mesh/mesh_topology.cpp:278-960); every block restricts its
     interior into a coarse buffer of the context (RestrictInterior,
     src/comm/refinement_ops.cpp:178-212); fine ghosts and interior
     boundary vertices take same-level slabs and finer neighbours'
     restricted slabs, coarse buffers take same-level coarse slabs and
     coarser neighbours' half-block slabs (CommChannel::PackAndSend /
     Unpack, src/comm/comm_channel.cpp:616-817, 1277-1472, ranges
     src/comm/index_utilities.cpp), ghosts start from zero and everything is
     divided by the node multiplicity (src/comm/node_multiplicity.cpp);
     then every block with a coarser neighbour fills its coarse ghosts at
     its physical faces (sommerfeld_faces | outflow_faces: the 5-point
     extrapolation over the coarse vertex ranges,
     ApplyPhysicalBCsOnCoarseLevel, src/comm/physical_bcs.cpp:687-876) and
     prolongates over each coarser neighbour's ProlongationRange
     (ProlongateBoundaries, src/comm/refinement_ops.cpp:143-176).  Fine
     ghosts beyond an open face are zero afterwards (for z4c_physical_bc).
     Z4C_EINVAL for a block list off such a tree, and for a block with a
     coarser neighbour at an open face that carries no physical boundary
     condition.  At most 22 fields per call on several spacings. */
int z4c_set_exchange(z4c_ctx* ctx, const double* lo, const double* hi, const int* periodic);
int z4c_exchange(z4c_ctx* ctx, int nvar, double* u, void* stream);

/* Transpose of z4c_exchange (no counterpart in the reference, which has no
   derivatives): ubar <- G^T ubar in place, nvar <= 22 fields.  On one
   spacing G is the same-level update (each destination vertex set to the
   mean of its lattice point's interior copies, every other vertex
   unchanged): a source vertex receives the cotangents of the destinations
   that read it (times 1/multiplicity), plus its own unless it is a
   destination itself; a ghost destination receives zero.  On several
   spacings G is the update across levels, P Phi_c E_g R (restriction, the
   gather, the coarse level's physical fill, prolongation), and the
   transpose runs their transposes in reverse order on ubar and a coarse
   cotangent; the prolongation's transpose counts each fine vertex it
   writes once however many boxes overlap there.  Deterministic (gathers in
   a fixed order).  The transpose data of the update across levels are
   built on the first call (device memory: about 22 doubles per source
   vertex and per coarse vertex).  Across levels the prolongation's
   transpose stages a block's full fine x range per workgroup, which limits
   blocks to about 140 cells along x (Z4C_EINVAL beyond). */
int z4c_exchange_vjp(z4c_ctx* ctx, int nvar, double* ubar, void* stream);

/* Transpose of z4c_physical_bc (no counterpart in the reference): the
   5-point extrapolation lines in reverse order, axes z, y, x. */
int z4c_physical_bc_vjp(z4c_ctx* ctx, int nvar, double* ubar, void* stream);

/* Physical ghost fill (PHY_BVAL of GR_Z4c: CommRegistry::ApplyPhysicalBCs,
   src/comm/comm_registry.cpp:516-527 -> comm::ApplyPhysicalBCs,
   src/comm/physical_bcs.cpp:506-676) on every face flagged in
   sommerfeld_faces | outflow_faces: the 5-point Lagrange extrapolation
   ExtrapolateOutflowBC (:187-290) that gr_sommerfeld and
   extrapolate_outflow faces of vertex-centred data get (:439-453), with the
   reference's ranges: it indexes the vertex arrays with the cell-centred
   is..ie = il..iu-1, so an outer face fills iu..iu+3 (the boundary vertex
   included, the outermost ghost not), and x2 / x3 faces span the x1 / x1
   and x2 ghosts.  Faces in order x1, x2, x3. */
int z4c_physical_bc(z4c_ctx* ctx, int nvar, double* u, void* stream);

/* Vertex-centred restriction and prolongation between refinement levels
   (ProlongOp::LagrangeUniform / RestrictOp::Injection of the Z4c fields,
   src/z4c/z4c.cpp:224-233; operators src/mesh/mesh_refinement_vc.cpp).
   Every block needs an even number of cells per axis.  A "coarse buffer"
   holds each block's coarse representation (MeshRefinement's coarse array,
   src/mesh/meshblock.cpp:611-640, 851-895): nx_d/2 + 9 vertices per axis,
   coarse interior 4 .. 4 + nx_d/2 at spacing 2 dx, coarse vertex 4 at the
   fine vertex il; the blocks sit back to back like a state buffer (block b
   at element offsets[b]*nvar, field n at + n*cn1*cn2*cn3).

   z4c_coarse_layout returns the per-block offsets (array of nblocks, may be
   NULL) and the elements one field needs (multiply by nvar). */
int z4c_coarse_layout(z4c_ctx* ctx, long long* offsets, long long* elems_per_var);

/* RestrictInterior (src/comm/refinement_ops.cpp:178-205) of every block:
   RestrictVertexCenteredValues (mesh_refinement_vc.cpp:56-100) injects the
   fine vertex at each coarse interior vertex.  Other coarse vertices are not
   written. */
int z4c_restrict(z4c_ctx* ctx, int nvar, const double* u, double* coarse, void* stream);

/* One call of ProlongateVertexCenteredValues (mesh_refinement_vc.cpp:283-772)
   on block `block` over the coarse index box [si, ei] x [sj, ej] x [sk, ek]. */
typedef struct {
  int block;
  int si, ei, sj, ej, sk, ek;
} z4c_prolong_box;

/* ProlongationRange (src/comm/refinement_ops.hpp:48-72, VC: pcng = 2) for a
   coarser neighbour of `block` at offset ox[3] (each -1, 0, 1) and the
   block's logical location lx[3] (only the parity matters): the box of
   ProlongateNeighbor (src/comm/refinement_ops.cpp:89-107).  Host only. */
int z4c_prolongation_range(z4c_ctx* ctx, int block, const int* ox, const long long* lx,
                           z4c_prolong_box* box);

/* z4c_set_prolongation keeps a list of boxes (host array; neighbour ghost
   boxes from z4c_prolongation_range, or a block's whole coarse interior as
   AMR refinement prolongates, src/comm/amr_registry.cpp:760-797); every box
   must keep its reads inside the coarse buffer and its writes inside the
   block (Z4C_EINVAL otherwise).  z4c_prolongate then applies all of them in
   one launch, fine vertices from the coarse buffer, in the reference's
   summation order; boxes that overlap write the same values there. */
int z4c_set_prolongation(z4c_ctx* ctx, const z4c_prolong_box* boxes, int nbox);
int z4c_prolongate(z4c_ctx* ctx, int nvar, const double* coarse, double* u, void* stream);

/* One bt_rk4 stage of the GR_Z4c task list
   (src/task_list/gr/task_list_gr_z4c.cpp:45-89, stage s = 1..4) on all
   blocks: k[s-1] = RHS(u_s) (+ Sommerfeld), u_{s+1} = u_n + sum a dt k
   (SumBT_ak / SumBT_bk), optional periodic ghost fill, AlgConstr,
   PrepareZ4cDerivatives.  `un` holds u at the start of the step; `u` is
   overwritten with the next stage state; k[4] are device pointers (host
   array).  Ghost update after SumBT, `periodic` = sum of: 1 single-block
   periodic fill (z4c_periodic_fill) or 2 the exchange set by
   z4c_set_exchange, and 4 the physical ghost fill (z4c_physical_bc);
   0 = none (ghosts stay u_n). */
int z4c_rk4_stage(z4c_ctx* ctx, const z4c_params* p, int stage, double dt,
                  double* u, const double* un, double* const* k,
                  const double* mat, int periodic, void* stream);

/* One whole bt_rk4 step of the GR_Z4c task list
   (src/task_list/gr/task_list_gr_z4c.cpp:45-89 over stages 1..4, tableau
   src/task_list/task_list_integrators_butcher.cpp:37-49) on all blocks, with
   ghost vertices held at u_n (no exchange) and Sommerfeld faces honoured.
   Per stage: the stencil pass of Z4cRHS, then one kernel that forms the
   curvature part, assembles the RHS and applies SumBT and AlgConstr at each
   interior vertex; the ghosts get AlgConstr(u_n) once.  The derivative
   arrays of PrepareZ4cDerivatives are not formed (nothing in the step reads
   them; z4c_adm_constraints prepares its own).  The state equals
   z4c_rk4_stage over stages 1..4 (same sums in the same order) without k
   registers: `work` is caller-owned device memory of 3 * 22 * elems_per_var
   doubles.  Without Sommerfeld faces every stage state is written over the
   interior of `u`, as the reference's stage update overwrites its state
   (work then holds u_n and the running sum); with them the stage states
   alternate between two work registers.  `u` holds u_n on entry and
   u_{n+1} on return. */
int z4c_rk4_step(z4c_ctx* ctx, const z4c_params* p, double dt, double* u,
                 double* work, const double* mat, void* stream);

/* z4c_rk4_step with the task list's ghost update after every stage instead
   of ghosts held at u_n: SumBT leaves u_n at the ghosts, then `ghosts` = 1
   (single-block periodic fill) or 2 (the exchange set by z4c_set_exchange),
   plus 4 (physical fill), then AlgConstr of the whole block -- the order and
   sums of z4c_rk4_stage over stages 1..4 (src/task_list/gr/
   task_list_gr_z4c.cpp:45-89) without its derivative preparation.  `work`
   as for z4c_rk4_step.  With the exchange on blocks of several spacings
   (z4c_set_exchange) every stage's update includes the restriction and the
   prolongation from coarser neighbours (GR_Z4c's PROLONG). */
int z4c_rk4_step_ghosts(z4c_ctx* ctx, const z4c_params* p, double dt, double* u,
                        double* work, const double* mat, int ghosts, void* stream);

/* z4c_rk4_step_ghosts (vacuum) replayed from a HIP graph (no counterpart in
   the reference, whose task list launches nothing): captured on a stream of
   the context on the first call and whenever u, work, dt, the parameters,
   `ghosts` or the exchange map change, then launched on `stream`.  The
   scratch of the ghost update is allocated before the capture, so the graph
   holds kernel nodes only.  Z4C_ESTATE while the kernel timer is on.
   Results equal z4c_rk4_step_ghosts'. */
int z4c_rk4_step_ghosts_graph(z4c_ctx* ctx, const z4c_params* p, double dt, double* u,
                              double* work, int ghosts, void* stream);

/* Low-storage (2S) integrator: per stage the weights delta, gamma_1,
   gamma_2, gamma_3 and beta of TaskLists::Integrators::LowStorage
   (src/task_list/task_list_integrators_low_storage.cpp:17-204, the
   IntegratorWeights of src/task_list/time_integrators.hpp). */
#define Z4C_LS_MAX_STAGES 5
typedef struct {
  int nstages;
  int pad_;
  double delta[Z4C_LS_MAX_STAGES];
  double gamma_1[Z4C_LS_MAX_STAGES];
  double gamma_2[Z4C_LS_MAX_STAGES];
  double gamma_3[Z4C_LS_MAX_STAGES];
  double beta[Z4C_LS_MAX_STAGES];
} z4c_ls_tableau;

/* The reference's tableau of <time>/integrator `name` (LowStorage
   constructor, task_list_integrators_low_storage.cpp:41-195): "vl2",
   "rk1", "rk2", "rk3", "rk4" (Ketcheson RK4()4[2S]; not bt_rk4).
   "ssprk5_4" and unknown names give Z4C_EINVAL, as GR_Z4c::StartupTaskList
   rejects ssprk5_4 (src/task_list/gr/task_list_gr_z4c.cpp:106-113).
   Host only. */
int z4c_ls_tableau_get(const char* name, z4c_ls_tableau* out);

/* One low-storage stage of the GR_Z4c task list (stage = 1..nstages;
   src/task_list/gr/task_list_gr_z4c.cpp:45-89): at stage 1 u1 is zeroed
   (StartupTaskList, :115-119); rhs = Z4cRHS(u) + Z4cBoundaryRHS; then
   IntegrateZ4c (:170-194): u1 = WeightedAve(u1, u, u2, {1, delta, 0}),
   u = WeightedAve(u, u1, u2, {gamma_1, gamma_2, gamma_3}),
   u += (beta dt) rhs (AddZ4cRHS); the ghost update selected by `ghosts`
   (as for z4c_rk4_stage); AlgConstr; PrepareZ4cDerivatives.  u, u1, rhs
   (and u2 when some gamma_3 != 0) are state buffers of 22 fields. */
int z4c_ls_stage(z4c_ctx* ctx, const z4c_params* p, const z4c_ls_tableau* t, int stage,
                 double dt, double* u, double* u1, const double* u2, double* rhs,
                 const double* mat, int ghosts, void* stream);

/* One whole low-storage step, fused per stage like z4c_rk4_step: the
   stencil pass, then one kernel forming the right-hand side, the two
   weighted averages, AddZ4cRHS and AlgConstr at each interior vertex (the
   sums and their order of z4c_ls_stage).  `ghosts` = 0 holds the ghosts
   out of the update (WeightedAve and AddZ4cRHS act on interior vertices),
   with AlgConstr applied to them after every stage; otherwise the ghost
   update of z4c_ls_stage (1 periodic fill or 2 exchange, plus 4 physical
   fill) and AlgConstr of the whole block after every stage.  The
   derivative arrays are not prepared.  `work` is caller-owned device
   memory of 2 * 22 * elems_per_var doubles (u1, and a second state buffer
   that Sommerfeld faces need).  2S tableaus only (every gamma_3 == 0, as
   for every integrator GR_Z4c accepts).  `u` holds u_n on entry and
   u_{n+1} on return. */
int z4c_ls_step(z4c_ctx* ctx, const z4c_params* p, const z4c_ls_tableau* t, double dt,
                double* u, double* work, const double* mat, int ghosts, void* stream);

/* z4c_rk4_step replayed from a HIP graph (no counterpart in the reference:
   its task list launches nothing).  The first call, and any call whose u,
   work, dt or parameters differ from the captured ones, captures the step's
   launches on a stream of the context; every call then launches the graph
   on `stream`.  Z4C_ESTATE while the kernel timer is on.  Results equal
   z4c_rk4_step's. */
int z4c_rk4_step_graph(z4c_ctx* ctx, const z4c_params* p, double dt, double* u,
                       double* work, void* stream);

/* Z4c constraint sums of the history output (src/outputs/history.cpp:178-205):
   out[8] = sum over blocks and over the reference's cell-index loop (vertices
   il..iu-1, jl..ju-1, kl..ku-1) of vol*{H^2, |M|, Mx^2, My^2, Mz^2, |Z|,
   Theta^2, |C|} with vol = dx*dy*dz, for cell centres inside r_max_con
   (the columns H-norm2 .. C-norm2 of the .hst file).  `con` comes from
   z4c_adm_constraints.  Waits for the stream; out is host memory. */
int z4c_history(z4c_ctx* ctx, const z4c_params* p, const double* u,
                const double* con, double* out, void* stream);

/* z4c_history without any allocation (no counterpart in the reference): for
   callers that cannot allocate device or host memory during the call (code
   under HIP graph capture, tight serving loops).  The per-tile partial sums
   live in a fixed device array of the library (8 MiB per GPU, held until the
   process ends) instead of a context allocation, and the columns are summed
   on the device in z4c_history's order, so `out` (8 doubles of DEVICE
   memory) equals z4c_history's result bit for bit.  Stream-ordered: does
   not wait.  At most 131072 interior tiles of 256 vertices (about 33.5 M
   interior vertices); beyond that Z4C_EINVAL.  Every call on one GPU uses
   the same array, so calls that overlap on one GPU (several streams or
   threads) race and must be serialised by the caller. */
int z4c_history_noalloc(z4c_ctx* ctx, const z4c_params* p, const double* u,
                        const double* con, double* out, void* stream);

/* Tile autotuner of the stencil pass (no counterpart in the reference,
   whose pencil loops have no tiling).  For every block size among the
   registered blocks, times the stencil pass of those blocks under each
   workgroup configuration (192, 256, 320, 384, 576 or 640 threads, each with the
   fewest x-y tiles that fit) on state `u`, `reps` launches each, and keeps
   the fastest.  Until it is called a block uses 256 threads, or 192 when
   its whole x-y plane fits one 192-thread tile.  Results do not depend on
   the configuration.
   Waits for the stream. */
int z4c_autotune(z4c_ctx* ctx, const z4c_params* p, const double* u, int reps, void* stream);
/* Workgroup size the stencil pass uses for `block`; setting it (block -1:
   all blocks) replaces the tuned or default choice. */
int z4c_stencil_threads(z4c_ctx* ctx, int block, int* threads);
int z4c_set_stencil_threads(z4c_ctx* ctx, int block, int threads);

/* Z4c::Z4cWeyl (src/z4c/calculate_weyl_scalars.cpp:44-628): weyl (2 fields
   per block, the state-buffer layout) receives Re psi4 and Im psi4 at the
   interior vertices, NaN elsewhere (as the reference fills u_weyl).  Inputs:
   the state u and its ADM variables adm (z4c_z4c_to_adm, 17 fields).  First
   prepares the conformal derivatives (PrepareZ4cDerivatives), which the
   reference has current when its Weyl task runs; the coordinate tetrad uses
   the vertex coordinates x0 + i*dx. */
int z4c_weyl(z4c_ctx* ctx, const z4c_params* p, const double* u, const double* adm,
             double* weyl, void* stream);

/* Passes of the kernel timer. */
enum {
  Z4C_PASS_STENCIL = 0,   /* derivative slots of Z4cRHS */
  Z4C_PASS_CURVATURE = 1, /* a separate Ricci / D_i D_j alpha pass: none in this build (both
                             z4c_rhs and z4c_rk4_step form it inside the assembly) */
  Z4C_PASS_ASSEMBLY = 2,  /* curvature part + RHS assembly (in z4c_rk4_step also SumBT and AlgConstr) */
  Z4C_PASS_GHOSTS = 3,    /* ghost pass of z4c_rk4_step */
  Z4C_PASS_PREPARE = 4,   /* PrepareZ4cDerivatives */
  Z4C_PASS_BT_SUM = 5,    /* SumBT */
  Z4C_PASS_ALGCONSTR = 6, /* AlgConstr */
  Z4C_NPASS = 7
};

/* Kernel timer (a profiling aid; no counterpart in the reference, whose
   timing is wall-clock per cycle, src/main.cpp).  While enabled, every pass
   launch is bracketed by HIP events on its own stream.  z4c_get_timing
   waits for the recorded events, returns the summed milliseconds and the
   launch count per pass (arrays of npass entries) and clears the record. */
int z4c_set_timing(z4c_ctx* ctx, int enable);
int z4c_get_timing(z4c_ctx* ctx, double* ms_total, long long* launches, int npass);

/* ------------------------------------------------------------------------
   Adjoints.  The reference has no derivative code; these are the
   vector-Jacobian products of its vacuum Z4cRHS (calculate_z4c_rhs.cpp:29-916,
   without Z4cBoundaryRHS, exactly what z4c_rhs computes) and of one fused
   bt_rk4 step (z4c_rk4_step), for sensitivities to the initial data and to
   the gauge parameters.  Parameter gradients per cotangent, in this order: */
enum { Z4C_PBAR_SHIFT_ETA = 0, Z4C_PBAR_DAMP_KAPPA1 = 1, Z4C_PBAR_DAMP_KAPPA2 = 2, Z4C_NPBAR = 3 };

/* Largest cotangent count of one adjoint call (the per-cotangent grid
   dimension of the adjoint kernels and the parameter-gradient reduction,
   nw * Z4C_NPBAR, stay within 65535). */
#define Z4C_MAX_NW 21845

/* VJP of z4c_rhs at state u for nw cotangents at once
   (1 <= nw <= Z4C_MAX_NW, Z4C_EINVAL otherwise):
   w holds nw registers of 22 * elems_per_var doubles (interior values are
   read); ubar receives nw registers, ubar_c = (d rhs/d u)^T w_c at every
   vertex of u, ghosts included; pbar (device, nw * Z4C_NPBAR doubles)
   receives d<w_c, rhs>/d(shift_eta, damp_kappa1, damp_kappa2).  Vacuum only;
   Z4C_EINVAL with cowling set.  The context allocates its scratch on the
   first call (282 + 162 nw doubles per interior vertex: the linearisation's
   derivatives and forward intermediates, and per cotangent the derivative
   cotangents and the pointwise adjoint's values between its two kernels). */
int z4c_rhs_vjp(z4c_ctx* ctx, const z4c_params* p, const double* u, int nw, const double* w,
                double* ubar, double* pbar, void* stream);

/* VJP of one z4c_rk4_step from u_n = u (unchanged): ubar_out holds nw
   cotangents of u_{n+1}, ubar_in receives those of u_n (every vertex) and
   pbar (nw * Z4C_NPBAR, device) the parameter gradients of <ubar_out, u_{n+1}>.
   The forward step is recomputed with its four stage right-hand sides and
   the stage states u_2..u_4 kept (stage checkpointing); the reverse sweep
   re-forms AlgConstr's input from u_n and the previous right-hand side and
   applies the AlgConstr adjoint at every stage.  Z4C_EINVAL with Sommerfeld
   faces or cowling.  Scratch: 7 + 3 nw registers of 22 * elems_per_var
   doubles on top of z4c_rhs_vjp's. */
int z4c_rk4_step_vjp(z4c_ctx* ctx, const z4c_params* p, double dt, const double* u, int nw,
                     const double* ubar_out, double* ubar_in, double* pbar, void* stream);

/* VJP of one z4c_rk4_step_ghosts with ghosts = 2 (the exchange after every
   stage, then AlgConstr) or 6 (the exchange, the physical fill, then
   AlgConstr), on one spacing or across levels; no counterpart in the
   reference.  Arguments as z4c_rk4_step_vjp.  The forward sweep keeps each
   stage's updated sum as AlgConstr's linearisation point; the reverse sweep
   applies AlgConstr's adjoint, z4c_physical_bc_vjp's and z4c_exchange_vjp's
   transposes and the stage sums' adjoint at every stage.  Z4C_EINVAL for
   other ghosts values, Sommerfeld faces, outflow faces with ghosts = 2, or
   cowling.  Scratch: 11 + 3 nw registers of 22 * elems_per_var doubles on
   top of z4c_rhs_vjp's and z4c_exchange_vjp's. */
int z4c_rk4_step_ghosts_vjp(z4c_ctx* ctx, const z4c_params* p, double dt, const double* u,
                            int nw, const double* ubar_out, double* ubar_in, double* pbar,
                            int ghosts, void* stream);

/* Volume integral of the squared Hamiltonian constraint of the Z4c
   right-hand side, Ht = R + 2/3 (Khat + 2 Theta)^2 - A~_ab A~^ab
   (calculate_z4c_rhs.cpp:606-624): J = sum over blocks and over the
   history's cell loop (vertices il..iu-1, jl..ju-1, kl..ku-1,
   src/outputs/history.cpp:178-205) of dx*dy*dz*Ht^2.  The objective of the
   adjoint sensitivity runs (the reference's H-norm2 column integrates the
   H of ADMConstraints, the same constraint discretised through the ADM
   variables).  Waits for the stream; J is host memory. */
int z4c_hamiltonian_norm(z4c_ctx* ctx, const z4c_params* p, const double* u, double* J,
                         void* stream);
/* ubar (one register of 22 * elems_per_var doubles) = dJ/du at every vertex. */
int z4c_hamiltonian_norm_vjp(z4c_ctx* ctx, const z4c_params* p, const double* u, double* ubar,
                             void* stream);

#ifdef __cplusplus
}
#endif
#endif /* Z4C_MI355X_H_ */
