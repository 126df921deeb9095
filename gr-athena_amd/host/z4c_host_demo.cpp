// z4c_host_demo.cpp -- a C++ caller of the C-ABI through the reference-side
// host (z4c_host.hpp): reads a block list, Z4c parameters and initial states,
// evaluates Z4cRHS of the initial state, then runs the GR_Z4c step loop, and
// writes both results.
//
// Usage: z4c_host_demo <input> <output> [<adjoint output>]
//
// Input (little-endian, packed):
//   char   magic[4] = "Z4CH"
//   int32  nblocks, nsteps, ghosts (GhostUpdate bits), fused (0/1)
//   char   integrator[16]            ("bt_rk4", "rk4", "rk3", ...)
//   double dt
//   double lo[3], hi[3]              domain of the exchange (used if ghosts & 2)
//   int32  periodic[3], pad
//   z4c_params                       (the header's struct, raw)
//   per block: int32 nx[3], int32 level, double xmin[3], double xmax[3],
//              int32 face[6] (0 block/periodic, 1 Sommerfeld, 2 outflow)
//   per block: 22 * nverts1 * nverts2 * nverts3 doubles, u(n,k,j,i)
// Output: per block the right-hand side of the initial state (22 fields,
// interior written, ghosts 0), then per block the state after nsteps steps.
// Adjoint output (optional): per block the RHS VJP at the initial state with
// that right-hand side as the cotangent, then its 3 parameter gradients;
// with the exchange (ghosts & 2), per block the exchange's transpose of the
// same VJP and then per block the VJP of one fused bt_rk4 step with the
// ghost update (cotangent: the initial state), and its 3 gradients.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "z4c_host.hpp"

// The header's structs as a C++ caller sees them: the layouts the Python
// binding and every other caller assume.
static_assert(sizeof(z4c_params) == 17 * 8 + 4 * 4, "z4c_params layout");
static_assert(sizeof(z4c_block) == 8 + 9 * 4 + 4 + 6 * 8 + 4 * 4, "z4c_block layout");
static_assert(offsetof(z4c_block, dx) == 48 && offsetof(z4c_block, level) == 96,
              "z4c_block field offsets");
static_assert(sizeof(z4c_ls_tableau) == 8 + 5 * Z4C_LS_MAX_STAGES * 8, "z4c_ls_tableau layout");
static_assert(sizeof(z4c_prolong_box) == 7 * 4, "z4c_prolong_box layout");

namespace {

template <class T>
T get(std::ifstream& in) {
  T v;
  in.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!in) throw std::runtime_error("input truncated");
  return v;
}

int run(const char* inpath, const char* outpath, const char* adjpath) {
  std::ifstream in(inpath, std::ios::binary);
  if (!in) throw std::runtime_error(std::string("cannot open ") + inpath);
  char magic[4];
  in.read(magic, 4);
  if (std::string(magic, 4) != "Z4CH") throw std::runtime_error("bad magic");
  const int nblocks = get<int32_t>(in), nsteps = get<int32_t>(in), ghosts = get<int32_t>(in),
            fused = get<int32_t>(in);
  char integ[17] = {0};
  in.read(integ, 16);
  const double dt = get<double>(in);
  double lo[3], hi[3];
  for (double& v : lo) v = get<double>(in);
  for (double& v : hi) v = get<double>(in);
  int periodic[3];
  for (int& v : periodic) v = get<int32_t>(in);
  (void)get<int32_t>(in);
  const z4c_params opt = get<z4c_params>(in);
  std::vector<z4c_host::MeshBlockInfo> blocks(nblocks);
  for (auto& b : blocks) {
    for (int& v : b.nx) v = get<int32_t>(in);
    b.level = get<int32_t>(in);
    for (double& v : b.xmin) v = get<double>(in);
    for (double& v : b.xmax) v = get<double>(in);
    for (auto& f : b.face) f = static_cast<z4c_host::Face>(get<int32_t>(in));
  }
  z4c_host::Z4cBatch batch(blocks, opt);
  std::vector<std::vector<double>> host(nblocks);
  for (int b = 0; b < nblocks; ++b) {
    host[b].resize((size_t)Z4C_NVAR * blocks[b].npts());
    in.read(reinterpret_cast<char*>(host[b].data()), host[b].size() * sizeof(double));
    if (!in) throw std::runtime_error("state truncated");
    batch.Upload(b, host[b].data(), Z4C_NVAR, batch.u.get());
  }
  if (ghosts & z4c_host::kExchange) batch.SetExchange(lo, hi, periodic);
  std::ofstream out(outpath, std::ios::binary);
  {
    z4c_host::DeviceBuffer rhs(Z4C_NVAR * batch.elems());
    batch.Z4cRHS(batch.u.get(), nullptr, rhs.get());
    batch.Z4cBoundaryRHS(batch.u.get(), rhs.get());
    batch.Synchronize();
    for (int b = 0; b < nblocks; ++b) {
      batch.Download(b, rhs.get(), Z4C_NVAR, host[b].data());
      out.write(reinterpret_cast<const char*>(host[b].data()), host[b].size() * sizeof(double));
    }
    if (adjpath) {
      std::ofstream adj(adjpath, std::ios::binary);
      z4c_host::DeviceBuffer ubar(Z4C_NVAR * batch.elems()), pbar(Z4C_NPBAR);
      std::vector<double> hp(Z4C_NPBAR), tmp;
      auto dump = [&](const double* d) {
        for (int b = 0; b < nblocks; ++b) {
          tmp.resize(host[b].size());
          batch.Download(b, d, Z4C_NVAR, tmp.data());
          adj.write(reinterpret_cast<const char*>(tmp.data()), tmp.size() * sizeof(double));
        }
      };
      auto dump_p = [&]() {
        z4c_host::HipCheck(hipMemcpy(hp.data(), pbar.get(), sizeof(double) * Z4C_NPBAR,
                                     hipMemcpyDeviceToHost),
                           "pbar");
        adj.write(reinterpret_cast<const char*>(hp.data()), hp.size() * sizeof(double));
      };
      batch.RhsVjp(batch.u.get(), 1, rhs.get(), ubar.get(), pbar.get());
      batch.Synchronize();
      dump(ubar.get());
      dump_p();
      if (ghosts & z4c_host::kExchange) {
        batch.ExchangeVjp(ubar.get());
        batch.Synchronize();
        dump(ubar.get());
        batch.StepGhostsVjp(dt, batch.u.get(), 1, batch.u.get(), ubar.get(), pbar.get(),
                            ghosts & (z4c_host::kExchange | z4c_host::kPhysical));
        batch.Synchronize();
        dump(ubar.get());
        dump_p();
      }
      if (!adj) throw std::runtime_error(std::string("cannot write ") + adjpath);
    }
  }
  for (int s = 0; s < nsteps; ++s) batch.Step(integ, dt, ghosts, fused != 0);
  batch.Synchronize();
  for (int b = 0; b < nblocks; ++b) {
    batch.Download(b, batch.u.get(), Z4C_NVAR, host[b].data());
    out.write(reinterpret_cast<const char*>(host[b].data()), host[b].size() * sizeof(double));
  }
  if (!out) throw std::runtime_error(std::string("cannot write ") + outpath);
  std::printf("z4c_host_demo: %d blocks, %d %s steps (%s), ghosts %d\n", nblocks, nsteps, integ,
              fused ? "fused" : "per stage", ghosts);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: %s <input> <output> [<adjoint output>]\n", argv[0]);
    return 2;
  }
  try {
    return run(argv[1], argv[2], argc == 4 ? argv[3] : nullptr);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "z4c_host_demo: %s\n", e.what());
    return 1;
  }
}
