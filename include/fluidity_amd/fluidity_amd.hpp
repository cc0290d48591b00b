// fluidity_amd.hpp -- the reference's split-solver template API over the C ABI.
//
// Header-only C++17.  The types keep the reference's names and template
// composition so that an instantiation written for the reference,
//
//   using solver_t = fluid::solver::SplitSolver<
//       fluid::solver::FaceFlux<
//           fluid::recon::MHReconstructor<fluid::limit::VanLeer<fluid::limit::cons_form_t>>,
//           fluid::flux::Hllc, fluid::material::IdealGas<double>>,
//       fluid::solver::BoundaryLoader<4>, fluid::Num<2>>;
//   solver_t solver{in};
//   solver.solve(in, out, mat, dtdh, setter);
//
// (solver/split_solver.hpp:37-126, solver/flux_solver.hpp:27-41,
// reconstruction/muscl_reconstructor.hpp:31-264, limiting/van_leer_limiter.hpp,
// flux_method/flux_hllc.hpp, material/ideal_gas.hpp) compiles against this
// header and runs the gfx950 kernels.  The compile-time choices become the
// runtime fl_scheme the C ABI takes.  Iterators are GridView values (a device
// pointer plus interior sizes) in place of the reference's MultidimIterator;
// solve() swaps the caller's in/out between sweeps as the reference does
// (solver/cuda/split_solver.cuh:75-78).  Errors throw fluid::amd::Error instead
// of the reference's print-and-terminate (utility/debug.hpp:42-67).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include <hip/hip_runtime_api.h>

#include "../fluidity_amd.h"

namespace fluid {

// Num<D> / Dimension<D> (dimension/dimension.hpp).
template <std::size_t D> struct Num : std::integral_constant<std::size_t, D> {};
template <std::size_t D> struct Dimension : std::integral_constant<std::size_t, D> {};
using dimx_t = Dimension<0>;
using dimy_t = Dimension<1>;
static constexpr dimx_t dim_x{};
static constexpr dimy_t dim_y{};

namespace amd {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(int status) {
  if (status != FL_OK)
    throw Error(status, std::string(fl_status_string(status)) + ": " + fl_last_error());
}

// Owns the fl_context of one device.
class Context {
 public:
  explicit Context(int device = 0) { check(fl_ctx_create(device, &ctx_)); }
  ~Context() { fl_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  fl_context* get() const { return ctx_; }
  static Context& instance(int device = 0) {
    static Context c(device);
    return c;
  }

 private:
  fl_context* ctx_ = nullptr;
};

template <typename T> constexpr int dtype_of();
template <> constexpr int dtype_of<double>() { return FL_F64; }
template <> constexpr int dtype_of<float>() { return FL_F32; }

// The iterator role of MultidimIterator<State<T, conservative, D>, DimInfo, gpu>
// (iterator/multidim_iterator.hpp:45-400): a device pointer to interior
// cells and the interior sizes, with the reference iterator's size(dim),
// stride(dim), num_dimensions(), offset(amount, dim) and shift(amount, dim).
// The entry points take an iterator at the grid's first cell, as the
// reference's simulator passes them; an offset iterator views the same
// memory from another cell.
template <typename T, std::size_t D, int Layout = FL_LAYOUT_AOS>
struct GridView {
  using value_t = T;
  static constexpr std::size_t dimensions = D;
  T* data = nullptr;
  int64_t n[3] = {1, 1, 1};

  GridView() = default;
  GridView(T* d, int64_t nx, int64_t ny = 1) : data(d) {
    n[0] = nx;
    n[1] = ny;
  }
  template <typename Dim> std::size_t size(Dim dim) const {
    return static_cast<std::size_t>(n[static_cast<std::size_t>(dim)]);
  }
  constexpr std::size_t num_dimensions() const { return D; }
  // Cells between neighbours in dimension dim (x fastest).
  template <typename Dim> std::size_t stride(Dim dim) const {
    std::size_t st = 1;
    for (std::size_t d = 0; d < static_cast<std::size_t>(dim); ++d) st *= size(d);
    return st;
  }
  template <typename Dim> GridView offset(int amount, Dim dim) const {
    GridView o = *this;
    o.shift(amount, dim);
    return o;
  }
  template <typename Dim> GridView& shift(int amount, Dim dim) {
    const int64_t cells = static_cast<int64_t>(amount) * static_cast<int64_t>(stride(dim));
    data += Layout == FL_LAYOUT_AOS ? cells * static_cast<int64_t>(2 + D) : cells;
    return *this;
  }
  T* get_ptr() const { return data; }
  std::size_t total_size() const { return size(0) * size(1) * size(2); }
  fl_field field() const {
    fl_field f{};
    f.data = data;
    f.dtype = dtype_of<T>();
    f.layout = Layout;
    f.ndim = static_cast<int32_t>(D);
    f.n[0] = n[0];
    f.n[1] = n[1];
    f.n[2] = n[2];
    return f;
  }
};

template <typename T, std::size_t D, int Layout = FL_LAYOUT_AOS>
using MultidimIterator = GridView<T, D, Layout>;

// Device memory of one array (the role of DeviceTensor<T, 1>,
// container/device_tensor.hpp): resize, begin/end as device pointers, and a
// 1D multi_iterator over it.
template <typename T>
class DeviceArray {
 public:
  struct iterator {
    using value_t = T;
    T* ptr = nullptr;
    T* get_ptr() const { return ptr; }
    friend std::ptrdiff_t operator-(const iterator& a, const iterator& b) { return a.ptr - b.ptr; }
  };
  DeviceArray() = default;
  explicit DeviceArray(std::size_t n) { resize(n); }
  ~DeviceArray() { release(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  void resize(std::size_t n) {
    if (n == size_) return;
    release();
    if (n && hipMalloc(reinterpret_cast<void**>(&data_), n * sizeof(T)) != hipSuccess)
      throw Error(FL_ERR_HIP, "hipMalloc of a device array failed");
    size_ = n;
  }
  std::size_t size() const { return size_; }
  T* data() const { return data_; }
  iterator begin() const { return iterator{data_}; }
  iterator end() const { return iterator{data_ + size_}; }
  GridView<T, 1> multi_iterator() const { return GridView<T, 1>(data_, (int64_t)size_); }

 private:
  void release() {
    if (data_) (void)hipFree(data_);
    data_ = nullptr;
    size_ = 0;
  }
  T* data_ = nullptr;
  std::size_t size_ = 0;
};

}  // namespace amd

// ---- limiting (limiting/limiter_traits.hpp, limiting/*.hpp) ----------------
namespace limit {
struct cons_form_t { static constexpr int form = FL_FORM_CONSERVATIVE; };
struct prim_form_t { static constexpr int form = FL_FORM_PRIMITIVE; };
template <typename Form> struct VanLeer {
  static constexpr int id = FL_LIMITER_VAN_LEER;
  static constexpr int form = Form::form;
  static constexpr std::size_t width = 4;  // LimiterTraits<VanLeer>::width
};
template <typename Form> struct Superbee {
  static constexpr int id = FL_LIMITER_SUPERBEE;
  static constexpr int form = Form::form;
  static constexpr std::size_t width = 4;  // LimiterTraits<Superbee>::width
};
}  // namespace limit

// ---- reconstruction (reconstruction/muscl_reconstructor.hpp) ---------------
namespace recon {
template <typename Limiter> struct MHReconstructor {
  using limiter_t = Limiter;
  static constexpr std::size_t width = Limiter::width;
};
}  // namespace recon

// ---- flux methods (flux_method/flux_hllc.hpp, flux_force.hpp) --------------
namespace flux {
struct Hllc { static constexpr int id = FL_FLUX_HLLC; };
struct Force { static constexpr int id = FL_FLUX_FORCE; };
struct LaxFriedrichs { static constexpr int id = FL_FLUX_LAX_FRIEDRICHS; };
struct Richtmyer { static constexpr int id = FL_FLUX_RICHTMYER; };
}  // namespace flux

// ---- materials (material/ideal_gas.hpp, material/stiff_fluid.hpp) ----------
namespace material {
template <typename T> struct IdealGas {
  using value_t = T;
  static constexpr int eos = FL_EOS_IDEAL;
  constexpr IdealGas() = default;
  constexpr explicit IdealGas(T adi) : adi_(adi) {}
  constexpr T adiabatic() const { return adi_; }
  constexpr T p_inf() const { return T(0); }

 private:
  T adi_ = T(1.4);
};
template <typename T> struct StiffFluid {
  using value_t = T;
  static constexpr int eos = FL_EOS_STIFF;
  constexpr StiffFluid() = default;
  constexpr explicit StiffFluid(T adi) : adi_(adi) {}
  constexpr StiffFluid(T adi, T p_inf) : adi_(adi), p_inf_(p_inf) {}
  constexpr T adiabatic() const { return adi_; }
  constexpr T p_inf() const { return p_inf_; }

 private:
  T adi_ = T(5.5);
  T p_inf_ = T(0.613);
};
}  // namespace material

// The material fields of an fl_scheme (eos, gamma, p_inf).
template <typename Mat>
fl_scheme material_scheme(const Mat& mat) {
  using M = std::decay_t<Mat>;
  fl_scheme s{};
  s.eos = M::eos;
  s.gamma = static_cast<double>(mat.adiabatic());
  s.p_inf = static_cast<double>(mat.p_inf());
  return s;
}

namespace solver {

// BoundaryKind / BoundaryIndex / BoundarySetter (solver/boundary_loader.hpp:29-124).
enum class BoundaryKind { transmissive = 0, reflective = 1, periodic = 2 };
enum class BoundaryIndex { first = 0, second = 1 };

// As in the reference, one kind applies to both ends of a dimension
// (its bit_index ignores the index: solver/boundary_loader.hpp:105-111).
struct BoundarySetter {
  template <typename Dim>
  void configure(Dim dim, BoundaryIndex, BoundaryKind kind) {
    kinds_[static_cast<std::size_t>(dim)] = static_cast<int32_t>(kind);
  }
  int32_t kind(std::size_t dim) const { return kinds_[dim]; }

 private:
  int32_t kinds_[3] = {0, 0, 0};
};

template <std::size_t Padding> struct BoundaryLoader {
  static constexpr std::size_t padding = Padding;
  static_assert(Padding == 4, "the MUSCL-Hancock stencil with chi flattening needs 4 ghosts");
};

template <typename ReconMethod, typename FluxMethod, typename Eos>
struct FaceFlux {
  using reconstructor_t = ReconMethod;
  using flux_method_t = FluxMethod;
  using eos_t = Eos;
  using value_t = typename Eos::value_t;
};

template <typename FluxSolver, typename Loader, typename Dims = Num<1>>
struct SplitSolver {
  using flux_solver_t = FluxSolver;
  using loader_t = Loader;
  using setter_t = BoundarySetter;
  using recon_t = typename FluxSolver::reconstructor_t;
  using limiter_t = typename recon_t::limiter_t;
  using eos_t = typename FluxSolver::eos_t;
  static constexpr std::size_t num_dimensions = Dims::value;
  static constexpr std::size_t padding = Loader::padding;

  SplitSolver() = default;
  template <typename It> explicit SplitSolver(It&&) {}
  template <typename It> void set_grid_sizes(It&&) {}

  static fl_scheme scheme(const eos_t& mat, const BoundarySetter& setter) {
    fl_scheme s{};
    s.limiter = limiter_t::id;
    s.form = limiter_t::form;
    s.flux = FluxSolver::flux_method_t::id;
    s.eos = eos_t::eos;
    s.gamma = static_cast<double>(mat.adiabatic());
    s.p_inf = static_cast<double>(mat.p_inf());
    for (int d = 0; d < 3; ++d) s.bc[d] = setter.kind(d);
    return s;
  }

  // SplitSolver::solve (solver/split_solver.hpp:105-115): all sweeps, x then y,
  // swapping the caller's in/out between sweeps.
  template <typename It, typename Mat, typename T>
  void solve(It& in, It& out, Mat&& mat, T dtdh, const BoundarySetter& setter,
             void* stream = nullptr) const {
    static_assert(std::decay_t<It>::dimensions == num_dimensions, "grid / solver dims differ");
    fl_field fi = in.field(), fo = out.field();
    const fl_scheme s = scheme(mat, setter);
    amd::check(fl_split_solve(amd::Context::instance().get(), &fi, &fo, &s,
                              static_cast<double>(dtdh), stream));
    if (fi.data != in.data) std::swap(in.data, out.data);
  }

  // One sweep in dimension Dim (solve_dim<S> -> SplitSolver::invoke).
  template <typename It, typename Mat, typename T, typename Dim>
  void sweep(const It& in, const It& out, Mat&& mat, T dtdh, const BoundarySetter& setter,
             Dim dim, void* stream = nullptr) const {
    fl_field fi = in.field(), fo = out.field();
    const fl_scheme s = scheme(mat, setter);
    amd::check(fl_split_sweep(amd::Context::instance().get(), &fi, &fo,
                              static_cast<int>(dim), &s, static_cast<double>(dtdh), stream));
  }
};

}  // namespace solver

// set_wavespeeds + max_element (simulator/wavespeed_initialization.hpp:34-53,
// algorithm/max_element.hpp:28-38) fused: the max wavespeed of a grid.
template <typename It, typename Mat>
double max_wavespeed(const It& in, const Mat& mat, void* stream = nullptr) {
  fl_field f = in.field();
  fl_scheme s = material_scheme(mat);
  double w = 0.0;
  amd::check(fl_max_wavespeed(amd::Context::instance().get(), &f, &s, &w, stream));
  return w;
}

namespace sim {

// set_wavespeeds (simulator/wavespeed_initialization.hpp:34-53): the FP64
// wavespeed of every cell of `states` into the array `wavespeeds` points at.
template <typename I1, typename I2, typename Mat>
void set_wavespeeds(I1&& states, I2&& wavespeeds, Mat&& mat) {
  fl_field f = states.field();
  fl_scheme s = material_scheme(mat);
  amd::check(fl_set_wavespeeds(amd::Context::instance().get(), &f, &s,
                               static_cast<double*>(wavespeeds.get_ptr()), nullptr));
}

}  // namespace sim

// max_element (algorithm/max_element.hpp:28-38) over a device array's range.
template <typename It, typename Return = typename std::decay_t<It>::value_t>
Return max_element(It&& begin, It&& end) {
  double m = 0.0;
  amd::check(fl_max_element(amd::Context::instance().get(), begin.get_ptr(),
                            static_cast<int64_t>(end - begin), &m, nullptr));
  return static_cast<Return>(m);
}

namespace sim {

// Domain (simulator/domain.hpp): the resolution the simulator uses.
struct Domain {
  explicit Domain(std::size_t dims = 1) : dims_(dims) {}
  void set_resolution(double r) { res_ = r; }
  double resolution() const { return res_; }
  std::size_t num_dimensions() const { return dims_; }

 private:
  std::size_t dims_;
  double res_ = 0.0;
};

// Parameters (simulator/parameters.hpp:30-99): cfl, run time, iterations and
// the host-side dt of the reference's loop, with the same expressions.
template <typename T>
struct Parameters {
  using value_t = std::decay_t<T>;
  using iter_t = std::size_t;
  static constexpr iter_t default_iters = static_cast<iter_t>(-1);

  value_t cfl = 0.9;
  value_t run_time = 0.0;
  value_t sim_time = 0.0;
  iter_t iters = 0;
  iter_t max_iters = default_iters;
  Domain domain;

  explicit Parameters(std::size_t num_dimensions) : domain(num_dimensions) {}
  void update_time_delta(value_t max_wavespeed) {
    dt_ = cfl * domain.resolution() / max_wavespeed;
  }
  void update_simulation_info() {
    run_time += dt_;
    iters += iter_t{1};
  }
  bool continue_simulation() const { return run_time < sim_time && iters < max_iters; }
  value_t dt() const { return dt_; }
  value_t dt_dh() const { return dt_ / domain.resolution(); }

 private:
  value_t dt_ = 0.0;
};

// SimulationData (simulator/simulation_data.hpp:40-140) on the device: input
// and output state grids, the wavespeed array, and swap(a, b) re-pointing the
// input at a and the output at b.  The grids are allocated here; upload()
// and download() move host data in the reference's AoS layout.
template <typename T, std::size_t D>
class SimulationData {
 public:
  using iter_t = amd::GridView<T, D>;

  SimulationData(int64_t nx, int64_t ny = 1)
      : nx_(nx), ny_(D > 1 ? ny : 1), a_(cells() * (2 + D)), b_(cells() * (2 + D)) {
    in_ = a_.data();
    out_ = b_.data();
  }
  std::size_t cells() const { return static_cast<std::size_t>(nx_ * ny_); }
  void initialize() { wavespeeds_.resize(cells()); }
  iter_t input_iterator() const { return iter_t(in_, nx_, ny_); }
  iter_t output_iterator() const { return iter_t(out_, nx_, ny_); }
  amd::GridView<double, D> wavespeed_iterator() const {
    return amd::GridView<double, D>(wavespeeds_.data(), nx_, ny_);
  }
  amd::DeviceArray<double>& wavespeeds() { return wavespeeds_; }
  const amd::DeviceArray<double>& wavespeeds() const { return wavespeeds_; }
  template <typename A, typename B>
  void swap(A&& a, B&& b) {
    in_ = a.get_ptr();
    out_ = b.get_ptr();
  }
  void upload(const T* host) {
    if (hipMemcpy(in_, host, cells() * (2 + D) * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw amd::Error(FL_ERR_HIP, "upload failed");
  }
  void download(T* host) const {
    if (hipMemcpy(host, in_, cells() * (2 + D) * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
      throw amd::Error(FL_ERR_HIP, "download failed");
  }

 private:
  int64_t nx_, ny_;
  amd::DeviceArray<T> a_, b_;
  amd::DeviceArray<double> wavespeeds_;
  T* in_ = nullptr;
  T* out_ = nullptr;
};

// The parameter surface of SimulatorImpl (simulator/simulator_impl.hpp:260-291)
// and its simulate() loop (:222-258), with dt kept on the device.
template <typename Solver>
class Simulator {
 public:
  using eos_t = typename Solver::eos_t;

  Simulator& configure_cfl(double cfl) { p_.cfl = cfl; return *this; }
  Simulator& configure_resolution(double res) { p_.dh = res; return *this; }
  Simulator& configure_sim_time(double t) { p_.sim_time = t; return *this; }
  Simulator& configure_max_iterations(int64_t it) { p_.max_iters = it; return *this; }
  Simulator& configure_material(const eos_t& m) { mat_ = m; return *this; }
  solver::BoundarySetter& setter() { return setter_; }

  // Advances `state` in place (`work` is a same-sized scratch grid).  As in
  // the reference, run time and iterations accumulate over calls
  // (simulator/parameters.hpp:43-76): a second call continues until the
  // accumulated values reach sim_time or max_iters.
  template <typename It>
  fl_sim_result simulate(It& state, It& work, void* stream = nullptr) {
    fl_field fs = state.field(), fw = work.field();
    const fl_scheme s = Solver::scheme(mat_, setter_);
    fl_sim_params p = p_;
    p.sim_time = p_.sim_time - run_time_;
    p.max_iters = p_.max_iters > iters_ ? p_.max_iters - iters_ : 0;
    fl_sim_result r{};
    amd::check(fl_simulate(ctx(), &fs, &fw, &s, &p, &r, stream));
    run_time_ += r.run_time;
    iters_ += r.iters;
    return r;
  }
  double run_time() const { return run_time_; }
  int64_t iterations() const { return iters_; }

  // print_results / write_results / write_results_separate_raw
  // (simulator/simulator_impl.hpp:343-420): the reference's ASCII formats of
  // `state`, stamped with the accumulated run time (sim_time for the raw
  // files) and placed with the configured resolution.
  template <typename It>
  void print_results(const It& state, void* stream = nullptr) const {
    write(state, nullptr, stream);
  }
  template <typename It>
  void write_results(const It& state, const std::string& prefix, void* stream = nullptr) const {
    write(state, (prefix + ".dat").c_str(), stream);
  }
  template <typename It>
  void write_results_separate_raw(const It& state, const std::string& prefix,
                                  void* stream = nullptr) const {
    fl_field fs = state.field();
    const fl_scheme s = Solver::scheme(mat_, setter_);
    amd::check(fl_write_results_separate_raw(amd::Context::instance().get(), &fs, &s,
                                             p_.sim_time, prefix.c_str(), stream));
  }

 private:
  template <typename It>
  void write(const It& state, const char* path, void* stream) const {
    fl_field fs = state.field();
    const fl_scheme s = Solver::scheme(mat_, setter_);
    amd::check(fl_write_results(amd::Context::instance().get(), &fs, &s, nullptr, run_time_,
                                p_.dh, path, stream));
  }

  fl_context* ctx() const { return amd::Context::instance().get(); }

  double run_time_ = 0.0;
  int64_t iters_ = 0;
  fl_sim_params p_{0.9, 0.0, 1e300, int64_t(1) << 40};
  eos_t mat_{};
  solver::BoundarySetter setter_{};
};

}  // namespace sim
}  // namespace fluid
