// C++ host-side mirror of the reference's Go package `sim` for the SPH step
// (sim/sph.go, sim/config-parser.go:104-149, sim/core.go:17-42), over the C
// ABI of include/sphugo.h.  Same names and argument meaning; where the
// reference panics (sim/sph.go:92-94,316-318) this throws sim::Panic.
//
//   sim::SphConfig conf = sim::MakeConfig();          // reference defaults
//   conf.HorPeriodicity = {0, 1}; conf.VertPeriodicity = {0, 1};
//   sim::Simulation s = sim::MakeSimulationFromConf(conf, particles);
//   s.Step();                                          // one Simulation.Step
//   double E = s.TotalEnergy();
//
// The particle slice keeps its input order (the device sorts internally);
// after every Step / CalculateForces the fields the reference's callers
// read (Pos, Vel, Rho, C, E, EDot, VDot, EPred, VPred, NNDists[0] = h) are
// refreshed from the device.
#pragma once

#include <cfloat>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sphugo.h"

namespace sim {

struct Panic : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(int rc) {
  if (rc != 0) throw Panic(sph_last_error());
}

// sim/linear-algebra.go:8-11
struct Vec2 {
  double X = 0, Y = 0;
};

// sim/sph.go:237-242; the device evaluates the kernel named by `id`
struct Kernel {
  int id;
  double FPrefactor, DFPrefactor;
};
// Monahan2D (sim/sph.go:258-276), Wendtland2D (:284-304): Go's folded constants
inline const Kernel Monahan2D{SPH_KERNEL_MONAGHAN, 0x1.5d3b3e3583243p+3, 0x1.5d3b3e3583243p+3};
inline const Kernel Wendtland2D{SPH_KERNEL_WENDLAND, 0x1.1d34a60108f72p+1, 0x1.1d34a60108f72p+2};

// sim/config-parser.go:104-109
struct Reflections {
  double L, R, U, D;
};

// The SphConfig fields the step reads (sim/config-parser.go:111-129), the
// neighbour count (NN_SIZE, sim/core.go:14) and the viscosity constants
// (sim/sph.go:379-383), plus the device-only choices (precision, order,
// CFL half step when DeltaTHalf <= 0).
struct SphConfig {
  int NSteps;
  double DeltaTHalf;
  double Gamma;
  double ParticleMass;
  Vec2 Acceleration;
  sim::Kernel Kernel;
  double HorPeriodicity[2];
  double VertPeriodicity[2];
  sim::Reflections Reflections;
  int NNSize = 32;
  double ViscAlpha = 0.75, ViscBeta = 1.5, ViscEta2 = 0.01;
  int Precision = SPH_PREC_FP64;
  bool ReferenceOrder = true;  // SPH_FLAG_BITWISE: neighbour sums in the reference's order
  double CFL = 0.3;
};

// MakeConfig, sim/config-parser.go:131-149
inline SphConfig MakeConfig() {
  SphConfig c{};
  c.Gamma = 1.66666;
  c.NSteps = 10000;
  c.DeltaTHalf = 0.001;
  c.ParticleMass = 1;
  c.Kernel = Monahan2D;
  c.VertPeriodicity[0] = -DBL_MAX;
  c.VertPeriodicity[1] = DBL_MAX;
  c.HorPeriodicity[0] = -DBL_MAX;
  c.HorPeriodicity[1] = DBL_MAX;
  c.Reflections = {-DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX};
  return c;
}

// Particle, sim/core.go:17-42 (the neighbour arrays reduce to NNDists[0])
struct Particle {
  Vec2 Pos, Vel;
  double Rho = 0, C = 0, E = 0, EDot = 0;
  Vec2 VDot;
  double EPred = 0;
  Vec2 VPred;
  double NNDist0 = 0;  // NNDists[0]: distance of the k-th neighbour (h)
};

class Simulation {
 public:
  SphConfig Config;
  std::vector<Particle> Particles;  // Root.Particles
  int CurrentStep = 0;

  Simulation(SphConfig conf, std::vector<Particle> ps)
      : Config(conf), Particles(std::move(ps)) {}
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;
  Simulation(Simulation&& o) noexcept
      : Config(o.Config), Particles(std::move(o.Particles)), CurrentStep(o.CurrentStep),
        dev_(o.dev_) {
    o.dev_ = nullptr;
  }
  Simulation& operator=(Simulation&& o) noexcept {
    if (this != &o) {
      sph_destroy(dev_);
      Config = o.Config;
      Particles = std::move(o.Particles);
      CurrentStep = o.CurrentStep;
      dev_ = o.dev_;
      o.dev_ = nullptr;
    }
    return *this;
  }
  ~Simulation() { sph_destroy(dev_); }

  // sim/sph.go:56-61
  void Run() {
    for (int s = 0; s < Config.NSteps; ++s) Step();
  }

  // sim/sph.go:64-198
  void Step() {
    if (Particles.empty()) throw Panic("int Run(): Simulation not initialized!");
    check(sph_step(handle(), 1));
    pull();
    CurrentStep += 1;
  }

  // n steps in one call, reading back once
  void StepN(int n) {
    if (Particles.empty()) throw Panic("int Run(): Simulation not initialized!");
    check(sph_step(handle(), n));
    pull();
    CurrentStep += n;
  }

  // sim/sph.go:403-435: kNN, density, sound speed from EPred, forces from
  // VPred, on the current device state
  void CalculateForces() {
    check(sph_calculate_forces(handle()));
    pull();
  }

  // sim/sph.go:441-463, over the slice in order, as the reference sums
  double TotalEnergy() const {
    double tot = 0.0;
    for (const Particle& p : Particles) tot += p.E;
    return tot;
  }
  double TotalDensity() const {
    double tot = 0.0;
    for (const Particle& p : Particles) tot += p.Rho;
    return tot;
  }
  double TotalMomentum() const {
    double tot = 0.0;
    for (const Particle& p : Particles) tot = __builtin_sqrt(p.Vel.X * p.Vel.X + p.Vel.Y * p.Vel.Y);
    return tot;
  }

  sph_sim* Device() { return handle(); }

  // The particles a source spawned (sim/sph.go:71-86: appended to
  // Root.Particles before the step); their IDs continue the slice's.
  void Append(const std::vector<Particle>& ps) {
    if (ps.empty()) return;
    const size_t m = ps.size();
    std::vector<double> buf(5 * m);
    for (size_t i = 0; i < m; ++i) {
      buf[i] = ps[i].Pos.X;
      buf[m + i] = ps[i].Pos.Y;
      buf[2 * m + i] = ps[i].Vel.X;
      buf[3 * m + i] = ps[i].Vel.Y;
      buf[4 * m + i] = ps[i].E;
    }
    check(sph_append(handle(), (int64_t)m, buf.data(), buf.data() + m, buf.data() + 2 * m,
                     buf.data() + 3 * m, buf.data() + 4 * m));
    Particles.insert(Particles.end(), ps.begin(), ps.end());
    pull();
  }

  // Checkpoint of the device state (sph_checkpoint_save): a byte image from
  // which Resume continues bit for bit.  The reference keeps no state on
  // disk; this has no counterpart there.
  std::vector<unsigned char> Checkpoint() {
    int64_t bytes = 0;
    check(sph_checkpoint_size(handle(), &bytes));
    std::vector<unsigned char> buf((size_t)bytes);
    check(sph_checkpoint_save(dev_, buf.data(), bytes));
    return buf;
  }

  // A simulation continued from a checkpoint: the device parameters come
  // from the image; conf supplies the host-side fields (NSteps).
  static Simulation Resume(const SphConfig& conf, const std::vector<unsigned char>& buf) {
    sph_sim* h = nullptr;
    check(sph_checkpoint_load(buf.data(), (int64_t)buf.size(), &h));
    Simulation s(conf, std::vector<Particle>((size_t)sph_num_particles(h)));
    s.dev_ = h;
    s.CurrentStep = (int)sph_current_step(h);
    s.pull();
    return s;
  }

 private:
  sph_sim* dev_ = nullptr;

  sph_params params() const {
    sph_params p;
    check(sph_default_params(&p));
    p.k = Config.NNSize;
    p.kernel = Config.Kernel.id;
    p.precision = Config.Precision;
    p.flags = Config.ReferenceOrder ? SPH_FLAG_BITWISE : 0u;
    p.gamma = Config.Gamma;
    p.mass = Config.ParticleMass;
    p.dt_half = Config.DeltaTHalf;
    p.cfl = Config.CFL;
    p.accel[0] = Config.Acceleration.X;
    p.accel[1] = Config.Acceleration.Y;
    for (int i = 0; i < 2; ++i) {
      p.hor[i] = Config.HorPeriodicity[i];
      p.vert[i] = Config.VertPeriodicity[i];
    }
    p.refl[0] = Config.Reflections.L;
    p.refl[1] = Config.Reflections.R;
    p.refl[2] = Config.Reflections.U;
    p.refl[3] = Config.Reflections.D;
    p.visc_alpha = Config.ViscAlpha;
    p.visc_beta = Config.ViscBeta;
    p.visc_eta2 = Config.ViscEta2;
    return p;
  }

  sph_sim* handle() {
    if (dev_) return dev_;
    const int64_t n = (int64_t)Particles.size();
    std::vector<double> buf(5 * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      const Particle& q = Particles[i];
      buf[i] = q.Pos.X;
      buf[n + i] = q.Pos.Y;
      buf[2 * n + i] = q.Vel.X;
      buf[3 * n + i] = q.Vel.Y;
      buf[4 * n + i] = q.E;
    }
    const sph_params p = params();
    check(sph_create(&p, n, buf.data(), buf.data() + n, buf.data() + 2 * n, buf.data() + 3 * n,
                     buf.data() + 4 * n, nullptr, &dev_));
    return dev_;
  }

  void pull() {
    const size_t n = Particles.size();
    std::vector<double> v(n);
    auto get = [&](int field, auto set) {
      check(sph_read_field(dev_, field, v.data()));
      for (size_t i = 0; i < n; ++i) set(Particles[i], v[i]);
    };
    get(SPH_X, [](Particle& q, double x) { q.Pos.X = x; });
    get(SPH_Y, [](Particle& q, double x) { q.Pos.Y = x; });
    get(SPH_VX, [](Particle& q, double x) { q.Vel.X = x; });
    get(SPH_VY, [](Particle& q, double x) { q.Vel.Y = x; });
    get(SPH_RHO, [](Particle& q, double x) { q.Rho = x; });
    get(SPH_C, [](Particle& q, double x) { q.C = x; });
    get(SPH_E, [](Particle& q, double x) { q.E = x; });
    get(SPH_EDOT, [](Particle& q, double x) { q.EDot = x; });
    get(SPH_AX, [](Particle& q, double x) { q.VDot.X = x; });
    get(SPH_AY, [](Particle& q, double x) { q.VDot.Y = x; });
    get(SPH_EPRED, [](Particle& q, double x) { q.EPred = x; });
    get(SPH_VPX, [](Particle& q, double x) { q.VPred.X = x; });
    get(SPH_VPY, [](Particle& q, double x) { q.VPred.Y = x; });
    get(SPH_H, [](Particle& q, double x) { q.NNDist0 = x; });
  }
};

// MakeSimulationFromConf (sim/sph.go:41-54) with the particles the
// config's Start spawners would produce handed in directly
inline Simulation MakeSimulationFromConf(const SphConfig& conf, std::vector<Particle> ps) {
  return Simulation(conf, std::move(ps));
}

}  // namespace sim
