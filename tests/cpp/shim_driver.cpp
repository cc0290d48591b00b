// Drives the reference-named template API of include/fluidity_amd/fluidity_amd.hpp
// the way the reference's SimulatorImpl drives SplitSolver, for the GPU tests:
//
//   shim_driver solve    <combo> <nd> <nx> <ny> <dtdh>  <in.bin> <out.bin>
//   shim_driver simulate <combo> <nd> <nx> <ny> <steps> <in.bin> <out.bin>
//   shim_driver write    <combo> <nd> <nx> <ny> <steps> <in.bin> <prefix>
//   shim_driver twice    <combo> <nd> <nx> <ny> <steps> <in.bin> <out.bin>
//   shim_driver impl     <combo> <nd> <nx> <ny> <steps> <in.bin> <out.bin>
//
// `twice` runs sim::Simulator::simulate with a limit of steps/2 iterations,
// raises the limit to `steps` and calls it again (run time and iterations
// accumulate across calls as in the reference).  `impl` runs RefSimulator
// below, whose simulate() body is SimulatorImpl::simulate's loop
// (simulator/simulator_impl.hpp:222-249) with its lines unchanged.
//
// `write` runs `steps` steps with sim_time 10, then writes <prefix>.dat
// (write_results) and <prefix>_<name>_10.000000.dat (write_results_separate_raw).
//
// Grids are raw little-endian float64 in the reference's AoS layout.  combo 0:
// MHReconstructor<VanLeer<cons>> + Hllc + IdealGas; combo 1: MHReconstructor<
// Superbee<prim>> + Force + IdealGas.  Prints run_time and iters for simulate.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/fluidity_amd/fluidity_amd.hpp"

using namespace fluid;

template <std::size_t D>
using vl_hllc_t = solver::SplitSolver<
    solver::FaceFlux<recon::MHReconstructor<limit::VanLeer<limit::cons_form_t>>, flux::Hllc,
                     material::IdealGas<double>>,
    solver::BoundaryLoader<4>, Num<D>>;
template <std::size_t D>
using sb_force_t = solver::SplitSolver<
    solver::FaceFlux<recon::MHReconstructor<limit::Superbee<limit::prim_form_t>>, flux::Force,
                     material::IdealGas<double>>,
    solver::BoundaryLoader<4>, Num<D>>;

static std::vector<double> read_file(const char* path, size_t n) {
  std::vector<double> v(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
    std::fprintf(stderr, "cannot read %s\n", path);
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

static void write_file(const char* path, const std::vector<double>& v) {
  FILE* f = std::fopen(path, "wb");
  if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
    std::fprintf(stderr, "cannot write %s\n", path);
    std::exit(2);
  }
  std::fclose(f);
}

// The members and loop of the reference's SimulatorImpl
// (simulator/simulator_impl.hpp:43-211, 222-258) over the shim's types; the
// loop body below is the reference's, line for line, minus its printing and
// timer.  It sits in fluid::sim as SimulatorImpl does, so the loop's
// unqualified names resolve as in the reference.
namespace fluid {
namespace sim {
template <typename Solver, std::size_t D>
class RefSimulator {
 public:
  using solver_t = Solver;
  using material_t = typename Solver::eos_t;

  RefSimulator(int64_t nx, int64_t ny) : _params(D), _data(nx, ny) {}

  void simulate() {
    auto solver = solver_t{_data.input_iterator()};
    auto mat    = material_t{};

    _data.initialize();

    auto cfl        = _params.cfl;
    auto wavespeeds = _data.wavespeed_iterator();
    (void)cfl;

    while (_params.continue_simulation())
    {
      auto in  = _data.input_iterator();
      auto out = _data.output_iterator();

      // Set the wavespeed data based on the updated state data from the previous
      // iteration, and then update sim time delta based on max wavespeed:
      set_wavespeeds(in, wavespeeds, mat);
      _params.update_time_delta(max_element(_data.wavespeeds().begin(),
                                            _data.wavespeeds().end()));

      solver.solve(in, out, mat, _params.dt_dh(), _setter);
      _params.update_simulation_info();
      _data.swap(out, in);
    }
  }

  Parameters<double> _params;
  SimulationData<double, D> _data;
  solver::BoundarySetter _setter;
};
}  // namespace sim
}  // namespace fluid

template <typename Solver, std::size_t D>
static int run(const std::string& mode, int64_t nx, int64_t ny, double arg, const char* in_path,
               const char* out_path) {
  const size_t cells = static_cast<size_t>(nx * ny);
  const size_t n = cells * (2 + D);
  std::vector<double> host = read_file(in_path, n);
  double *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(double)) != hipSuccess ||
      hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 3;
  }
  (void)hipMemcpy(a, host.data(), n * sizeof(double), hipMemcpyHostToDevice);
  amd::GridView<double, D> in(a, nx, ny), out(b, nx, ny);
  material::IdealGas<double> mat(1.4);
  solver::BoundarySetter setter;
  setter.configure(dim_x, solver::BoundaryIndex::first, solver::BoundaryKind::transmissive);
  try {
    if (mode == "impl") {
      (void)hipFree(a);
      (void)hipFree(b);
      sim::RefSimulator<Solver, D> sim(nx, ny);
      sim._params.cfl = 0.9;
      sim._params.sim_time = 1e300;
      sim._params.max_iters = static_cast<std::size_t>(arg);
      sim._params.domain.set_resolution(1.0 / nx);
      sim._data.upload(host.data());
      sim.simulate();
      sim._data.download(host.data());
      std::printf("run_time %.17g iters %llu\n", sim._params.run_time,
                  (unsigned long long)sim._params.iters);
      write_file(out_path, host);
      return 0;
    }
    if (mode == "twice") {
      sim::Simulator<Solver> simr;
      const int64_t steps = static_cast<int64_t>(arg);
      simr.configure_cfl(0.9).configure_resolution(1.0 / nx).configure_max_iterations(steps / 2);
      simr.configure_material(mat);
      simr.simulate(in, out);
      simr.configure_max_iterations(steps);
      const fl_sim_result r = simr.simulate(in, out);
      std::printf("run_time %.17g iters %lld last %lld\n", simr.run_time(),
                  (long long)simr.iterations(), (long long)r.iters);
      (void)hipMemcpy(host.data(), in.data, n * sizeof(double), hipMemcpyDeviceToHost);
      write_file(out_path, host);
      (void)hipFree(a);
      (void)hipFree(b);
      return 0;
    }
    if (mode == "solve") {
      Solver solver{in};
      solver.solve(in, out, mat, arg, setter);  // result in `out` afterwards
      (void)hipDeviceSynchronize();
      (void)hipMemcpy(host.data(), out.data, n * sizeof(double), hipMemcpyDeviceToHost);
    } else {
      sim::Simulator<Solver> simr;
      simr.configure_cfl(0.9).configure_resolution(1.0 / nx).configure_max_iterations(
          static_cast<int64_t>(arg));
      simr.configure_material(mat);
      if (mode == "write") simr.configure_sim_time(10.0);
      const fl_sim_result r = simr.simulate(in, out);
      std::printf("run_time %.17g iters %lld\n", r.run_time, (long long)r.iters);
      if (mode == "write") {
        simr.write_results(in, out_path);
        simr.write_results_separate_raw(in, out_path);
        (void)hipFree(a);
        (void)hipFree(b);
        return 0;
      }
      (void)hipMemcpy(host.data(), in.data, n * sizeof(double), hipMemcpyDeviceToHost);
    }
  } catch (const amd::Error& e) {
    std::fprintf(stderr, "fluidity_amd error: %s\n", e.what());
    return 4;
  }
  write_file(out_path, host);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 9) {
    std::fprintf(stderr, "usage: %s solve|simulate|write|twice|impl combo nd nx ny arg in.bin out\n",
                 argv[0]);
    return 1;
  }
  const std::string mode = argv[1];
  const int combo = std::atoi(argv[2]);
  const int nd = std::atoi(argv[3]);
  const int64_t nx = std::atoll(argv[4]), ny = std::atoll(argv[5]);
  const double arg = std::atof(argv[6]);
  if (nd == 1 && combo == 0) return run<vl_hllc_t<1>, 1>(mode, nx, 1, arg, argv[7], argv[8]);
  if (nd == 1 && combo == 1) return run<sb_force_t<1>, 1>(mode, nx, 1, arg, argv[7], argv[8]);
  if (nd == 2 && combo == 0) return run<vl_hllc_t<2>, 2>(mode, nx, ny, arg, argv[7], argv[8]);
  if (nd == 2 && combo == 1) return run<sb_force_t<2>, 2>(mode, nx, ny, arg, argv[7], argv[8]);
  std::fprintf(stderr, "unsupported combo/nd\n");
  return 1;
}
