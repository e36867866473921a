// C++ drop-in check: drives the GPU step through the reference-shaped C++ layer
// (cil::hip::CellColony, Session, EulerHIP as an MDIntegrator) and compares the result
// with the CPU oracle on the same input.  Exit codes: 0 ok, 1 mismatch, 3 no GPU device.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../include/cil_hip_md.hpp"
#include "../../oracle/cil_oracle.h"

typedef void (*MDIntegrator)(int& ts);  // MD.cuh:450

int main() {
  const int n = 150, nsteps = 200;
  const double L = 18.0;
  cil_params p;
  std::memset(&p, 0, sizeof p);
  p.abi_version = CIL_ABI_VERSION;
  p.dim = 2;
  p.nspecies = 1;
  cil_species& sp = p.species[0];
  sp.nelem = 2;
  for (int e = 0; e < 2; ++e) {
    sp.sigma[e] = 1.0;
    sp.zeta[e] = 1.0;
    sp.eps_core[e] = 1.0;
    sp.sig_well[e] = 1.0;
  }
  sp.motility[1] = 0.1;
  sp.spring = CIL_SPRING_HARMONIC;
  sp.k_harm = 10.0;
  sp.cross_adhesion = 1.0;
  sp.cycle_p0 = 1.0;
  sp.kT = 0.01;
  for (int i = 0; i < CIL_MAX_TYPES; ++i)
    for (int j = 0; j < CIL_MAX_TYPES; ++j) {
      p.law[i][j] = (i < 2 && j < 2) ? CIL_LAW_SC : CIL_LAW_NONE;
      p.rcut2[i][j] = 1.0;
    }
  p.dt = 1e-4;
  p.noise = CIL_NOISE_FDT;
  p.seed = 7;
  p.precision = CIL_PREC_FP64;
  cil_box box;
  std::memset(&box, 0, sizeof box);
  box.length[0] = box.length[1] = L;
  box.periodic[0] = box.periodic[1] = 1;

  cil::hip::CellColony gpu(2, 2, {n}), cpu(2, 2, {n}), b1(2, 2, {n}), b2(2, 2, {n}), l32(2, 2, {n});
  std::mt19937_64 rng(5);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  for (int c = 0; c < n; ++c) {
    int a = gpu.addCell(0);
    cpu.addCell(0);
    double cx = U(rng) * L, cy = U(rng) * L, th = U(rng) * 6.283185307179586;
    double ux = std::cos(th), uy = std::sin(th);
    double xb[2] = {std::fmod(cx - 0.25 * ux + L, L), std::fmod(cy - 0.25 * uy + L, L)};
    double xf[2] = {std::fmod(cx + 0.25 * ux + L, L), std::fmod(cy + 0.25 * uy + L, L)};
    for (int d = 0; d < 2; ++d) {
      gpu.pos(0, a, 0, d) = cpu.pos(0, a, 0, d) = xb[d];
      gpu.pos(0, a, 1, d) = cpu.pos(0, a, 1, d) = xf[d];
    }
    for (int e = 0; e < 2; ++e) gpu.sigma(0, a, e) = cpu.sigma(0, a, e) = 0.9 + 0.2 * U(rng);
    gpu.species(0).rest_len[a] = cpu.species(0).rest_len[a] = 0.5;
    gpu.species(0).state[a] = cpu.species(0).state[a] = 1;
  }
  b1 = gpu;  // batch members: the same colony (seed 7) and a second one with seed 8
  b2 = gpu;
  l32 = gpu;  // FP32 forces with neighbour lists
  try {
    cil::hip::Session session(p, box, gpu);
    session.upload();
    session.calculForcesPressure();  // CIL.cu:1096
    cil::hip::bind(&session);
    MDIntegrator mdStep = cil::hip::EulerHIP;
    int ts = 0;
    for (int k = 0; k < nsteps; ++k) mdStep(ts);
    session.download();
    // oracle on the host copy
    or_ws* w = or_create(&p, &box);
    or_set_traversal(w, 1);
    auto v = cpu.views();
    int tso = 0;
    if (or_forces(w, v.data(), 1, 0) || or_step(w, v.data(), 1, nsteps, &tso)) {
      std::printf("oracle failed: %s\n", or_last_error(w));
      return 1;
    }
    or_destroy(w);
    double err = 0.0;
    for (int c = 0; c < n; ++c)
      for (int e = 0; e < 2; ++e)
        for (int d = 0; d < 2; ++d) {
          double x = gpu.pos(0, c, e, d) - cpu.pos(0, c, e, d);
          x -= L * std::round(x / L);
          err = std::fmax(err, std::fabs(x));
        }
    std::printf("ts=%d max position difference %.3e\n", ts, err);
    // the same colony inside a two-colony batch must match the single session bit for bit
    cil_params p2 = p;
    p2.seed = 8;
    cil::hip::Session s1(p, box, b1), s2(p2, box, b2);
    s1.upload();
    s2.upload();
    s1.calculForcesPressure();
    s2.calculForcesPressure();
    std::vector<int> tsb = {0, 0};
    {
      cil::hip::Batch batch({&s1, &s2});
      batch.steps(nsteps, tsb);
    }
    s1.download();
    s2.download();
    bool same = true, differs = false;
    for (int c = 0; c < n; ++c)
      for (int e = 0; e < 2; ++e)
        for (int d = 0; d < 2; ++d) {
          same = same && b1.pos(0, c, e, d) == gpu.pos(0, c, e, d);
          differs = differs || b2.pos(0, c, e, d) != gpu.pos(0, c, e, d);
        }
    std::printf("batch: colony 0 %s the single session, colony 1 (other seed) %s\n",
                same ? "equals" : "DIFFERS FROM", differs ? "differs" : "DOES NOT DIFFER");
    // FP32 forces over neighbour lists through the same layer, against the FP64 oracle
    cil_params p32 = p;
    p32.precision = CIL_PREC_FP32;
    cil::hip::Session s3(p32, box, l32);
    s3.upload();
    s3.setNeighbourList(true, 0.3, 160);
    s3.calculForcesPressure();
    int ts3 = 0;
    s3.steps(nsteps, ts3);
    s3.download();
    double err32 = 0.0;
    for (int c = 0; c < n; ++c)
      for (int e = 0; e < 2; ++e)
        for (int d = 0; d < 2; ++d) {
          double x = l32.pos(0, c, e, d) - cpu.pos(0, c, e, d);
          x -= L * std::round(x / L);
          err32 = std::fmax(err32, std::fabs(x));
        }
    std::printf("FP32 neighbour-list session: max position difference %.3e\n", err32);
    // the reference's seeded run from its own JSON file: two frames through the C++ layer,
    // written in the reference's GSD format from the device, against the oracle's frames
    bool run_ok = true;
    if (const char* cfg = std::getenv("CIL_CHECK_CONFIG")) {
      auto rc = cil::hip::RunConfig::read(cfg);
      cil::hip::CellColony rg(rc.params.dim, rc.params.species[0].nelem, {64}), ro(rc.params.dim, 1, {64});
      rc.initColony(rg);
      ro = rg;
      cil::hip::Session s4(rc.params, rc.box, rg);
      s4.upload();
      s4.calculForcesPressure();
      or_ws* w4 = or_create(&rc.params, &rc.box);
      or_set_traversal(w4, 1);
      auto v4 = ro.views();
      int tsg = 0, tso4 = 0;
      or_forces(w4, v4.data(), 1, 0);
      {
        cil::hip::GsdFile gsd(std::string(std::getenv("CIL_CHECK_OUT") ? std::getenv("CIL_CHECK_OUT") : "/tmp") +
                              "/check.gsd");
        for (int f = 0; f < 2; ++f) {
          s4.steps(rc.run.gts, tsg);
          gsd.frame(s4, &rc.run, tsg);
          or_step(w4, v4.data(), 1, rc.run.gts, &tso4);
        }
      }
      ro.pull_counts(v4);
      s4.download();
      double e4 = 0.0;
      for (int c = 0; c < rg.getNumCells(0); ++c)
        for (int d = 0; d < 2; ++d) e4 = std::fmax(e4, std::fabs(rg.pos(0, c, 0, d) - ro.pos(0, c, 0, d)));
      or_destroy(w4);
      run_ok = tsg == tso4 && rg.getNumCells(0) == ro.getNumCells(0) && e4 <= 1e-9;
      std::printf("seeded run from its configuration: %d cells after %d steps, %.3e from the oracle\n",
                  rg.getNumCells(0), tsg, e4);
    }
    return (ts == nsteps && err <= 1e-9 && same && differs && tsb[0] == nsteps && tsb[1] == nsteps &&
            ts3 == nsteps && err32 <= 1e-5 && run_ok)
               ? 0
               : 1;
  } catch (const std::runtime_error& e) {
    if (std::strstr(e.what(), "status -2")) {
      std::printf("no GPU device: %s\n", e.what());
      return 3;
    }
    std::printf("error: %s\n", e.what());
    return 1;
  }
}
