// cil_hip_md.hpp — C++ host mirror of the reference's step interfaces over the C ABI.
//
// The reference drives the step through process-global state (CIL.cu:13-23): a CellColony
// (Colony.hpp:16-209), the MDIntegrator function pointer `mdStep` (MD.cuh:450-452) called
// GTS times per frame (CIL.cu:1106), calculForcesPressure at start-up (CIL.cu:1096) and
// checkPositions once per frame (CIL.cu:1107).  This header gives the same shapes:
//
//   cil::hip::CellColony     per-species arrays in the reference layout, addCell/removeCell
//   cil::hip::Session        owns one device context (upload, step, download)
//   cil::hip::EulerHIP       `void (*)(int& ts)`-compatible step on the bound session
//                            (StochasticHeunHIP: the same seam for STOCHASTIC_HEUN runs)
//   cil::hip::Batch          many sessions stepped by one launch per kernel
//   cil::hip::RunConfig      a reference JSON run file read into parameters, box and run
//                            settings, and its initial colony (readMD/readSpecs/..., the INIT
//                            block and initializeCellStatus on the reference's stream)
//   cil::hip::GsdFile, H5File  frames in the reference's GSD and HDF5 formats from a session
//
// Errors surface as std::runtime_error, the reference's run_assert convention
// (Utils.hpp:99-109).  Host arrays are never cached by the device side.
#pragma once

#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cil_hip.h"

namespace cil {
namespace hip {

struct SpeciesArrays {
  int ncells = 0;
  std::vector<double> pos, frc, vel, sigma;                       // (cap, nelem[, dim])
  std::vector<double> sigma_t, pressure, theta, rho, ang_vel, rest_len;  // (cap)
  std::vector<int> state, t0, tswitch, tpressure;                 // (cap)
};

// Per-species SoA store with the reference's row-major layout pos(cell, elem, d)
// (Array.h:557-600) and its append / swap-with-last semantics (Colony.cpp:355-399).
class CellColony {
 public:
  CellColony(int dim, int nelem, const std::vector<int>& capacity) : dim_(dim), nelem_(nelem) {
    for (int cap : capacity) {
      SpeciesArrays s;
      size_t c = (size_t)cap;
      s.pos.assign(c * nelem * dim, 0.0);
      s.frc.assign(c * nelem * dim, 0.0);
      s.vel.assign(c * nelem * dim, 0.0);
      s.sigma.assign(c * nelem, 1.0);
      for (auto* v : {&s.sigma_t, &s.pressure, &s.theta, &s.rho, &s.ang_vel, &s.rest_len}) v->assign(c, 0.0);
      s.state.assign(c, 2);
      s.t0.assign(c, 0);
      s.tswitch.assign(c, -1);
      s.tpressure.assign(c, 0);
      spec_.push_back(std::move(s));
      cap_.push_back(cap);
    }
  }
  int getCellTypes() const { return (int)spec_.size(); }
  int getNumCells(int s) const { return spec_[s].ncells; }
  int getTotalCells() const {
    int n = 0;
    for (auto& s : spec_) n += s.ncells;
    return n;
  }
  int capacity(int s) const { return cap_[s]; }
  int dim() const { return dim_; }
  int nelem() const { return nelem_; }
  double& pos(int s, int c, int e, int d) { return spec_[s].pos[((size_t)c * nelem_ + e) * dim_ + d]; }
  double& frc(int s, int c, int e, int d) { return spec_[s].frc[((size_t)c * nelem_ + e) * dim_ + d]; }
  double& sigma(int s, int c, int e) { return spec_[s].sigma[(size_t)c * nelem_ + e]; }
  SpeciesArrays& species(int s) { return spec_[s]; }

  int addCell(int s) {  // Colony.cpp:355-359: append, returns the new index
    if (spec_[s].ncells >= cap_[s]) throw std::runtime_error("CellColony::addCell: capacity exhausted");
    return spec_[s].ncells++;
  }

  std::vector<cil_species_view> views() {
    std::vector<cil_species_view> v(spec_.size());
    for (size_t k = 0; k < spec_.size(); ++k) {
      SpeciesArrays& s = spec_[k];
      cil_species_view& w = v[k];
      w.ncells = s.ncells;
      w.capacity = cap_[k];
      w.pos = s.pos.data();
      w.frc = s.frc.data();
      w.vel = s.vel.data();
      w.sigma = s.sigma.data();
      w.sigma_t = s.sigma_t.data();
      w.pressure = s.pressure.data();
      w.state = s.state.data();
      w.t0 = s.t0.data();
      w.tswitch = s.tswitch.data();
      w.tpressure = s.tpressure.data();
      w.theta = s.theta.data();
      w.rho = s.rho.data();
      w.ang_vel = s.ang_vel.data();
      w.rest_len = s.rest_len.data();
    }
    return v;
  }
  void pull_counts(const std::vector<cil_species_view>& v) {
    for (size_t k = 0; k < spec_.size(); ++k) spec_[k].ncells = v[k].ncells;
  }

 private:
  int dim_, nelem_;
  std::vector<SpeciesArrays> spec_;
  std::vector<int> cap_;
};

inline void run_assert(int rc, cil_ctx* ctx, const char* what) {
  if (rc != CIL_OK)
    throw std::runtime_error(std::string(what) + ": " + (ctx ? cil_hip_last_error(ctx) : "") + " (status " +
                             std::to_string(rc) + ")");
}

// One device context bound to a host colony.  Device state is resident between calls;
// the host arrays are only touched by upload() and download().
class Session {
 public:
  Session(const cil_params& p, const cil_box& box, CellColony& col) : col_(col) {
    std::vector<int> cap(CIL_MAX_SPECIES, 0);
    for (int s = 0; s < col.getCellTypes(); ++s) cap[s] = col.capacity(s);
    run_assert(cil_hip_create(&p, &box, cap.data(), &ctx_), nullptr, "cil_hip_create");
  }
  ~Session() { cil_hip_destroy(ctx_); }
  Session(const Session&) = delete;
  Session& operator=(const Session&) = delete;

  void upload() {
    auto v = col_.views();
    run_assert(cil_hip_upload(ctx_, v.data(), (int)v.size()), ctx_, "cil_hip_upload");
  }
  // calculForcesPressure (CIL.cu:999-1007), as called once before the frame loop (CIL.cu:1096)
  void calculForcesPressure() { run_assert(cil_hip_forces(ctx_, nullptr, nullptr), ctx_, "cil_hip_forces"); }
  // `gts` Euler steps (the frame's inner loop, CIL.cu:1106)
  void steps(int gts, int& ts) { run_assert(cil_hip_step(ctx_, gts, &ts), ctx_, "cil_hip_step"); }
  // checkPositions (CIL.cu:945-966); returns the number of drained cells
  int checkPositions() {
    int removed = 0;
    run_assert(cil_hip_check_positions(ctx_, &removed), ctx_, "checkPositions");
    return removed;
  }
  void download() {
    auto v = col_.views();
    std::vector<int> need(v.size());
    run_assert(cil_hip_download(ctx_, v.data(), (int)v.size(), need.data()), ctx_, "cil_hip_download");
    col_.pull_counts(v);
  }
  // Bins per cut-off and pair-kernel block (project addition; cil_hip_set_tuning)
  void setTuning(int bins_per_cutoff, int block) {
    run_assert(cil_hip_set_tuning(ctx_, bins_per_cutoff, block), ctx_, "cil_hip_set_tuning");
  }
  // Neighbour-list mode for FP32 forces (project addition; cil_hip_set_neighbour_list)
  void setNeighbourList(bool on, double skin = 0.3, int max_neighbours = 256) {
    run_assert(cil_hip_set_neighbour_list(ctx_, on ? 1 : 0, skin, max_neighbours), ctx_,
               "cil_hip_set_neighbour_list");
  }
  // Two-level lists over the list above (project addition; cil_hip_set_neighbour_list_inner)
  void setInnerList(double skin_in) {
    run_assert(cil_hip_set_neighbour_list_inner(ctx_, skin_in), ctx_, "cil_hip_set_neighbour_list_inner");
  }
  // Division records for the HDF5 /div_app group (project addition; cil_hip_set_division_log)
  void setDivisionLog(bool on) {
    run_assert(cil_hip_set_division_log(ctx_, on ? 1 : 0), ctx_, "cil_hip_set_division_log");
  }
  CellColony& colony() { return col_; }
  cil_ctx* ctx() { return ctx_; }

 private:
  CellColony& col_;
  cil_ctx* ctx_ = nullptr;
};

// A reference run file (project addition over cil_hip_config_read / cil_hip_config_init).
struct RunConfig {
  cil_params params{};
  cil_box box{};
  cil_run run{};
  std::string text;

  static RunConfig read(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("RunConfig: cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    RunConfig c;
    c.text = ss.str();
    char err[512] = {0};
    if (cil_hip_config_read(c.text.c_str(), &c.params, &c.box, &c.run, err, sizeof err) != CIL_OK)
      throw std::runtime_error(std::string("RunConfig: ") + err);
    return c;
  }
  // The INIT cells into `col` (capacity at least run.ncells[s] per species); sets
  // params.rng_offset to the draws initializeCellStatus consumed.
  void initColony(CellColony& col) {
    auto v = col.views();
    char err[512] = {0};
    if (cil_hip_config_init(text.c_str(), &params, v.data(), (int)v.size(), err, sizeof err) != CIL_OK)
      throw std::runtime_error(std::string("RunConfig::initColony: ") + err);
    col.pull_counts(v);
  }
};

// outputGSDFrame's file (project addition over cil_hip_gsd_*).
class GsdFile {
 public:
  explicit GsdFile(const std::string& path) {
    if (cil_hip_gsd_create(path.c_str(), "CIL", "hoomd", (1u << 16) | 1u, &g_) != CIL_OK)
      throw std::runtime_error("GsdFile: cannot create " + path);
  }
  ~GsdFile() { cil_hip_gsd_close(g_); }
  GsdFile(const GsdFile&) = delete;
  GsdFile& operator=(const GsdFile&) = delete;
  void frame(Session& s, const cil_run* run, int ts) {
    run_assert(cil_hip_gsd_write_frame_ctx(s.ctx(), g_, run, ts), s.ctx(), "cil_hip_gsd_write_frame_ctx");
  }

 private:
  cil_gsd* g_ = nullptr;
};

// initHDF5Output / outputHDF5Frame's file (project addition over cil_hip_h5_*).
class H5File {
 public:
  H5File(const std::string& path, const cil_params& p, const cil_box& box, const cil_run* run) {
    if (cil_hip_h5_create(path.c_str(), &p, &box, run, &f_) != CIL_OK)
      throw std::runtime_error("H5File: cannot create " + path + " (is libhdf5.so loadable?)");
  }
  ~H5File() { cil_hip_h5_close(f_); }
  H5File(const H5File&) = delete;
  H5File& operator=(const H5File&) = delete;
  void frame(Session& s, const cil_run* run, int id, int ts, double cpu_time = 0.0) {
    run_assert(cil_hip_h5_write_frame_ctx(s.ctx(), f_, run, id, ts, cpu_time), s.ctx(), "cil_hip_h5_write_frame_ctx");
  }

 private:
  cil_h5* f_ = nullptr;
};

// Many sessions advanced together, one launch per kernel for all colonies
// (cil_hip_batch_*).  The sessions must outlive the batch.
class Batch {
 public:
  explicit Batch(const std::vector<Session*>& sessions) {
    std::vector<cil_ctx*> c;
    for (Session* s : sessions) c.push_back(s->ctx());
    run_assert(cil_hip_batch_create(c.data(), (int)c.size(), &b_), nullptr, "cil_hip_batch_create");
  }
  ~Batch() { cil_hip_batch_destroy(b_); }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  // `gts` Euler steps of every colony; ts[c] is colony c's step counter
  void steps(int gts, std::vector<int>& ts) {
    int rc = cil_hip_batch_step(b_, gts, ts.data());
    if (rc != CIL_OK)
      throw std::runtime_error(std::string("cil_hip_batch_step: ") + cil_hip_batch_last_error(b_) + " (status " +
                               std::to_string(rc) + ")");
  }
  cil_batch* handle() { return b_; }

 private:
  cil_batch* b_ = nullptr;
};

// The MDIntegrator-shaped entry point: bind a session once, then assign
// `mdStep = cil::hip::EulerHIP;` in place of Euler (Input.cpp:526-532).
inline Session*& bound_session() {
  static Session* s = nullptr;
  return s;
}
inline void bind(Session* s) { bound_session() = s; }
inline void EulerHIP(int& ts) {
  Session* s = bound_session();
  if (!s) throw std::runtime_error("EulerHIP: no session bound");
  s->steps(1, ts);
}
// StochasticHeun (CIL.cu:1028-1049): the session's parameters select the integrator
// (cil_params.integrator = CIL_INT_HEUN), so the seam is the same one step.
inline void StochasticHeunHIP(int& ts) { EulerHIP(ts); }

}  // namespace hip
}  // namespace cil
