// z4c_host.hpp -- reference-side C++ host of the MI355X Z4c path.
//
// What a GR-Athena++ rank holds to drive the C-ABI of include/z4c_mi355x.h:
//
//  * MeshBlockInfo: the MeshBlock fields the path reads -- block_size
//    (RegionSize, src/athena.hpp:88-96: cell counts, the coordinates of the
//    first and last vertex), the logical level, and the per-face boundary
//    kind -- with the vertex-centred index ranges of
//    src/mesh/meshblock.cpp:869-889 (ivs = NGHOST, ive = ivs + nx,
//    nverts = nx + 1 + 2 NGHOST);
//  * Z4cBatch: one device context for all blocks of the rank, the Z4c
//    registers in the reference's AthenaArray order u(n,k,j,i) per block,
//    the Z4c method names of src/z4c/z4c.hpp:448-610 acting on every block
//    in one call, and the GR_Z4c step (src/task_list/gr/task_list_gr_z4c.cpp:
//    45-89) over the Butcher bt_rk4 or a low-storage integrator
//    (<time>/integrator), per stage or fused.
//
// Errors from the C-ABI become std::runtime_error (the reference's
// ATHENA_ERROR path), with the library's message.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "z4c_mi355x.h"

namespace z4c_host {

inline void Check(int rc, const char* what) {
  if (rc != Z4C_OK) throw std::runtime_error(std::string(what) + ": " + z4c_last_error());
}
inline void HipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Per-face boundary kinds of a block (BoundaryFlag in the reference): a
// neighbour block or a periodic image, a Sommerfeld (gr_sommerfeld) face, or
// an extrapolated outflow face.
enum class Face : int { block = 0, sommerfeld = 1, outflow = 2 };

struct MeshBlockInfo {
  static constexpr int NGHOST = 4;
  int nx[3];         // block_size.nx1..nx3 (cells)
  double xmin[3];    // block_size.x1min..x3min: coordinate of vertex ivs
  double xmax[3];    // block_size.x1max..x3max: coordinate of vertex ive
  int level = 0;     // loc.level
  Face face[6] = {Face::block, Face::block, Face::block,
                  Face::block, Face::block, Face::block};  // ix1, ox1, ix2, ox2, ix3, ox3

  int nverts(int d) const { return nx[d] + 1 + 2 * NGHOST; }
  int ivs(int /*d*/) const { return NGHOST; }
  int ive(int d) const { return NGHOST + nx[d]; }
  double dx(int d) const { return (xmax[d] - xmin[d]) / nx[d]; }
  long long npts() const { return (long long)nverts(0) * nverts(1) * nverts(2); }

  // z4c_block of this MeshBlock at element offset `offset`
  z4c_block Describe(long long offset) const {
    z4c_block d{};
    d.offset = offset;
    d.nn1 = nverts(0);
    d.nn2 = nverts(1);
    d.nn3 = nverts(2);
    d.il = ivs(0); d.iu = ive(0);
    d.jl = ivs(1); d.ju = ive(1);
    d.kl = ivs(2); d.ku = ive(2);
    for (int a = 0; a < 3; ++a) {
      d.dx[a] = dx(a);
      d.x0[a] = xmin[a] - NGHOST * dx(a);
    }
    d.level = level;
    for (int f = 0; f < 6; ++f) {
      if (face[f] == Face::sommerfeld) d.sommerfeld_faces |= 1 << f;
      if (face[f] == Face::outflow) d.outflow_faces |= 1 << f;
    }
    return d;
  }
};

// Device memory of `n` doubles, owned.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) : n_(n) {
    HipCheck(hipMalloc(&p_, sizeof(double) * n), "hipMalloc");
    HipCheck(hipMemset(p_, 0, sizeof(double) * n), "hipMemset");
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~DeviceBuffer() {
    if (p_) (void)hipFree(p_);
  }
  double* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  double* p_ = nullptr;
  size_t n_ = 0;
};

// Ghost update after each stage (SEND/RECV/SETB_Z4C and PHY_BVAL of GR_Z4c).
enum GhostUpdate : int { kGhostsHeld = 0, kPeriodicSelf = 1, kExchange = 2, kPhysical = 4 };

class Z4cBatch {
 public:
  Z4cBatch(const std::vector<MeshBlockInfo>& blocks, const z4c_params& opt, int device = 0)
      : blocks_(blocks), opt_(opt) {
    HipCheck(hipSetDevice(device), "hipSetDevice");
    Check(z4c_ctx_create(&ctx_, device), "z4c_ctx_create");
    std::vector<z4c_block> d;
    long long off = 0;
    for (const MeshBlockInfo& b : blocks_) {
      d.push_back(b.Describe(off));
      offsets_.push_back(off);
      off += b.npts();
    }
    Check(z4c_set_blocks(ctx_, d.data(), (int)d.size(), &elems_), "z4c_set_blocks");
    if (elems_ != off) throw std::runtime_error("z4c_set_blocks: unexpected layout");
    u = DeviceBuffer(Z4C_NVAR * elems_);
  }
  ~Z4cBatch() {
    if (ctx_) z4c_ctx_destroy(ctx_);
  }
  Z4cBatch(const Z4cBatch&) = delete;
  Z4cBatch& operator=(const Z4cBatch&) = delete;

  long long elems() const { return elems_; }
  int nblocks() const { return (int)blocks_.size(); }
  z4c_params& opt() { return opt_; }

  // Block b's nvar fields between a host AthenaArray (n,k,j,i) and a device
  // buffer of the batch layout.
  void Upload(int b, const double* host, int nvar, double* dev) const {
    const size_t n = (size_t)nvar * blocks_[b].npts();
    HipCheck(hipMemcpy(dev + offsets_[b] * nvar, host, n * sizeof(double), hipMemcpyHostToDevice),
             "upload");
  }
  void Download(int b, const double* dev, int nvar, double* host) const {
    const size_t n = (size_t)nvar * blocks_[b].npts();
    HipCheck(hipMemcpy(host, dev + offsets_[b] * nvar, n * sizeof(double), hipMemcpyDeviceToHost),
             "download");
  }

  // -- the Z4c class surface (src/z4c/z4c.hpp:448-610), all blocks --------
  void Z4cRHS(const double* uin, const double* mat, double* rhs) {
    Check(z4c_rhs(ctx_, &opt_, uin, mat, rhs, nullptr), "Z4cRHS");
  }
  void Z4cBoundaryRHS(const double* uin, double* rhs) {
    Check(z4c_boundary_rhs(ctx_, &opt_, uin, rhs, nullptr), "Z4cBoundaryRHS");
  }
  void AlgConstr(double* uio) { Check(z4c_algconstr(ctx_, &opt_, uio, nullptr), "AlgConstr"); }
  void ADMToZ4c(const double* adm, double* uout) {
    Check(z4c_adm_to_z4c(ctx_, &opt_, adm, uout, nullptr, nullptr), "ADMToZ4c");
  }
  void Z4cToADM(const double* uin, double* adm) {
    Check(z4c_z4c_to_adm(ctx_, &opt_, uin, adm, nullptr, nullptr), "Z4cToADM");
  }
  void ADMConstraints(double* con, const double* adm, const double* mat, const double* uin) {
    Check(z4c_adm_constraints(ctx_, &opt_, uin, adm, mat, con, nullptr), "ADMConstraints");
  }
  double NewBlockTimeStep(double cfl) {
    double dt = 0.0;
    Check(z4c_new_dt(ctx_, cfl, &dt), "NewBlockTimeStep");
    return dt;
  }
  // the same-level exchange over the domain [lo, hi] (periodic axes wrap)
  void SetExchange(const double lo[3], const double hi[3], const int periodic[3]) {
    Check(z4c_set_exchange(ctx_, lo, hi, periodic), "z4c_set_exchange");
  }
  void Exchange(double* uio) { Check(z4c_exchange(ctx_, Z4C_NVAR, uio, nullptr), "exchange"); }

  // -- adjoints (no counterpart in the reference) ----------------------------
  // ubar (nw registers) = (d rhs / d u)^T w at uin; pbar (nw * Z4C_NPBAR,
  // device) the gauge-parameter gradients
  void RhsVjp(const double* uin, int nw, const double* w, double* ubar, double* pbar) {
    Check(z4c_rhs_vjp(ctx_, &opt_, uin, nw, w, ubar, pbar, nullptr), "z4c_rhs_vjp");
  }
  // ubar <- (the ghost update set by SetExchange)^T ubar
  void ExchangeVjp(double* ubar) {
    Check(z4c_exchange_vjp(ctx_, Z4C_NVAR, ubar, nullptr), "z4c_exchange_vjp");
  }
  // VJP of the fused bt_rk4 step with the ghost update (ghosts 2 or 6)
  void StepGhostsVjp(double dt, const double* uin, int nw, const double* ubar_out,
                     double* ubar_in, double* pbar, int ghosts) {
    Check(z4c_rk4_step_ghosts_vjp(ctx_, &opt_, dt, uin, nw, ubar_out, ubar_in, pbar, ghosts,
                                  nullptr),
          "z4c_rk4_step_ghosts_vjp");
  }

  // -- GR_Z4c: one step of <time>/integrator on u ----------------------------
  // "bt_rk4": Butcher stages (SumBT_ak / SumBT_bk); "vl2", "rk1", "rk2",
  // "rk3", "rk4": low-storage stages (IntegrateZ4c).  `fused`: one call per
  // step with the stage kernel fused (no derivative preparation) instead of
  // the task list's calls per stage.
  void Step(const std::string& integrator, double dt, int ghosts, bool fused) {
    if (integrator == "bt_rk4") {
      if (fused) {
        Ensure(work_, 3);
        Check(z4c_rk4_step_ghosts(ctx_, &opt_, dt, u.get(), work_.get(), nullptr, ghosts, nullptr),
              "z4c_rk4_step_ghosts");
        return;
      }
      Ensure(un_, 1);
      Ensure(k_, 4);
      // PutScratchBT_u at stage 1 (task_list_integrators_butcher.cpp:542-571)
      HipCheck(hipMemcpy(un_.get(), u.get(), sizeof(double) * Z4C_NVAR * elems_,
                         hipMemcpyDeviceToDevice), "u_n");
      double* k[4];
      for (int s = 0; s < 4; ++s) k[s] = k_.get() + (size_t)s * Z4C_NVAR * elems_;
      for (int s = 1; s <= 4; ++s)
        Check(z4c_rk4_stage(ctx_, &opt_, s, dt, u.get(), un_.get(), k, nullptr, ghosts, nullptr),
              "z4c_rk4_stage");
      return;
    }
    z4c_ls_tableau t;
    Check(z4c_ls_tableau_get(integrator.c_str(), &t), "z4c_ls_tableau_get");
    if (fused) {
      Ensure(work_, 3);
      Check(z4c_ls_step(ctx_, &opt_, &t, dt, u.get(), work_.get(), nullptr, ghosts, nullptr),
            "z4c_ls_step");
      return;
    }
    Ensure(un_, 1);  // u1
    Ensure(k_, 1);   // rhs
    for (int s = 1; s <= t.nstages; ++s)
      Check(z4c_ls_stage(ctx_, &opt_, &t, s, dt, u.get(), un_.get(), nullptr, k_.get(), nullptr,
                         ghosts, nullptr),
            "z4c_ls_stage");
  }

  void Synchronize() { HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

  DeviceBuffer u;  // storage.u of every block

 private:
  void Ensure(DeviceBuffer& b, int nreg) {
    const size_t n = (size_t)nreg * Z4C_NVAR * elems_;
    if (b.size() < n) b = DeviceBuffer(n);
  }
  std::vector<MeshBlockInfo> blocks_;
  std::vector<long long> offsets_;
  z4c_params opt_;
  z4c_ctx* ctx_ = nullptr;
  long long elems_ = 0;
  DeviceBuffer work_, un_, k_;
};

}  // namespace z4c_host
