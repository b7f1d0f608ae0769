// Runs the small uniform case through the C++ mirror of package `sim`
// (sim.hpp) and prints the reference's totals as one JSON line.
//
//   sim_demo [n=2200] [steps=10] [seed=12345678] [k=32] [resume_at=-1]
//
// With resume_at >= 0 the run is checkpointed after resume_at steps and
// continued from the checkpoint in a new Simulation (same output expected).
//
// Initial conditions: SplitMix64 counter stream (value i = mix(seed +
// (i+1) * 0x9E3779B97F4A7C15)), top 53 bits scaled by 2^-53; x takes values
// 1..n, y values n+1..2n; v = 0, e = 1, m = 1/n, gamma = 5/3, periodic unit
// box, alpha = 1, beta = 2, CFL 0.3 half step.
#include <cstdio>
#include <cstdlib>
#include <cstdint>

#include "sim.hpp"

static uint64_t splitmix(uint64_t seed, uint64_t i) {
  uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main(int argc, char** argv) {
  const int64_t n = argc > 1 ? std::atoll(argv[1]) : 2200;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 10;
  const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 12345678ull;
  const int k = argc > 4 ? std::atoi(argv[4]) : 32;
  const int resume_at = argc > 5 ? std::atoi(argv[5]) : -1;
  std::vector<sim::Particle> ps((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    ps[i].Pos.X = (double)(splitmix(seed, (uint64_t)i + 1) >> 11) * 0x1p-53;
    ps[i].Pos.Y = (double)(splitmix(seed, (uint64_t)(n + i) + 1) >> 11) * 0x1p-53;
    ps[i].E = 1.0;
  }
  sim::SphConfig conf = sim::MakeConfig();
  conf.NSteps = steps;
  conf.NNSize = k;
  conf.Gamma = 5.0 / 3.0;
  conf.ParticleMass = 1.0 / (double)n;
  conf.DeltaTHalf = 0.0;  // CFL half step
  conf.HorPeriodicity[0] = 0;
  conf.HorPeriodicity[1] = 1;
  conf.VertPeriodicity[0] = 0;
  conf.VertPeriodicity[1] = 1;
  conf.ViscAlpha = 1.0;
  conf.ViscBeta = 2.0;
  try {
    sim::Simulation s = sim::MakeSimulationFromConf(conf, std::move(ps));
    if (resume_at >= 0 && resume_at <= steps) {
      for (int i = 0; i < resume_at; ++i) s.Step();
      const std::vector<unsigned char> ckpt = s.Checkpoint();
      s = sim::Simulation::Resume(conf, ckpt);
      for (int i = resume_at; i < steps; ++i) s.Step();
    } else {
      s.Run();
    }
    double sx = 0;
    for (const sim::Particle& q : s.Particles) sx += q.Pos.X;
    std::printf("{\"n\": %lld, \"steps\": %d, \"step\": %d, \"TotalEnergy\": %.17g, "
                "\"TotalDensity\": %.17g, \"TotalMomentum\": %.17g, \"sum_x\": %.17g, "
                "\"h_first\": %.17g}\n",
                (long long)n, steps, s.CurrentStep, s.TotalEnergy(), s.TotalDensity(),
                s.TotalMomentum(), sx, s.Particles[0].NNDist0);
  } catch (const sim::Panic& e) {
    std::fprintf(stderr, "panic: %s\n", e.what());
    return 2;
  }
  return 0;
}
