// fl_oracle_io.cpp -- CPU restatement of the reference's initial-condition
// setters and result writers (test infrastructure only: the checker of the
// library's fl_from_primitive / fl_set_states / fl_write_results /
// fl_write_results_separate_raw).  Written with the C++ streams and
// manipulators the reference uses, so that the library's printf-based
// formatting is checked against the standard library's own.
//
// States are the oracle's AoS conservative layout [rho, E, m...] per cell,
// x fastest.
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

extern "C" {
#include "fl_oracle.h"
}

namespace {

const char* const kNames[] = {"rho", "p", "v_x", "v_y"};

// IdealGas::eos (material/ideal_gas.hpp:60-63) / StiffFluid::eos
// (material/stiff_fluid.hpp:77-80) of a primitive state.
double eos_of(const double* W, const orc_scheme* s) {
  if (s->eos == ORC_EOS_STIFF || s->eos == ORC_EOS_STIFF_CORRECTED) return (W[1] + s->gamma * s->p_inf) / ((s->gamma - 1.0) * W[0]);
  return W[1] / ((s->gamma - 1.0) * W[0]);
}

// stream_output_ascii, single material (simulator/simulator_impl.hpp:435-508)
// and two-material (simulator/multimaterial_simulator.hpp:882-971), with the
// reference's offset bookkeeping.
void stream_output(std::ostream& stream, const double* U, const int64_t* n, int nd,
                   const orc_scheme* s, const double* phi, double run_time,
                   double resolution) {
  const int S = 2 + nd;
  const std::string comment("# ");
  stream << comment << "t = " << std::to_string(run_time) << "\n";
  int column = 1;
  for (int d = 0; d < nd; ++d)
    stream << comment << "Column " << column++ << ": " << char(120 + d) << " coordinate\n";
  if (phi)
    for (int d = 0; d < nd; ++d)
      stream << comment << "Column " << column++ << ": " << char(120 + d)
             << " ls coordinate\n";
  for (int k = 0; k < S; ++k) stream << comment << "Column " << column++ << ": " << kNames[k] << "\n";
  if (phi)
    stream << comment << "Column " << column << ": " << "internal energy (e)\n";
  else
    stream << comment << "Column " << column << " : " << "internal energy (e)\n";

  auto append_element = [&stream](double element) {
    stream << std::setw(12) << std::left << std::fixed << std::showpoint
           << std::setprecision(8) << element << " ";
  };

  std::vector<std::size_t> size(nd), offsets(nd, 0);
  for (int d = 0; d < nd; ++d) size[d] = (std::size_t)n[d];
  std::size_t row = 0;  // the iterator's shift in y
  do {
    for (int i = 0; i < nd - 1; ++i)
      if (offsets[i] == size[i]) offsets[i] = 0;
    for (std::size_t x = 0; x < size[0]; ++x) {
      const std::size_t cell = row * size[0] + x;
      const double* u = U + cell * S;
      double W[4];
      orc_cons_to_prim(u, nd, 0, s, W);
      for (int d = 0; d < nd; ++d) append_element((offsets[d] + 0.5) * resolution);
      if (phi)
        for (int d = 0; d < nd; ++d) append_element(phi[cell]);
      for (int k = 0; k < S; ++k) append_element(W[k]);
      append_element(eos_of(W, s));
      stream << "\n";
      offsets[0]++;
    }
    for (int i = 0; i < nd - 1; ++i)
      if (offsets[i] == size[i]) {
        offsets[i + 1]++;
        row++;
      }
  } while (offsets[nd - 1] != size[nd - 1]);
}

}  // namespace

extern "C" {

int orc_write_ascii(const double* U, const int64_t* n, int nd, const orc_scheme* s,
                    const double* phi, double run_time, double resolution, const char* path) {
  std::ofstream f(path, std::fstream::trunc);
  if (!f) return 1;
  stream_output(f, U, n, nd, s, phi, run_time, resolution);
  f.close();
  return f ? 0 : 1;
}

// write_results_separate_raw / write_blob (simulator/simulator_impl.hpp:360-420).
int orc_write_separate_raw(const double* U, const int64_t* n, int nd, const orc_scheme* s,
                           double sim_time, const char* prefix) {
  const int S = 2 + nd;
  std::vector<std::ofstream> streams;
  for (int k = 0; k < S; ++k)
    streams.emplace_back(std::string(prefix) + "_" + kNames[k] + "_" +
                         std::to_string(sim_time) + ".dat");
  for (auto& st : streams)
    if (!st) return 1;
  const int64_t rows = nd > 1 ? n[1] : 1;
  for (int64_t y = 0; y < rows; ++y) {
    for (int64_t x = 0; x < n[0]; ++x) {
      double W[4];
      orc_cons_to_prim(U + (y * n[0] + x) * S, nd, 0, s, W);
      for (int k = 0; k < S; ++k) streams[k] << W[k] << " ";
    }
    for (auto& st : streams) st << "\n";
  }
  if (nd == 1)
    for (auto& st : streams) st << "\n";
  for (auto& st : streams) st.close();
  return 0;
}

// set_states_impl (state/cuda/state_initialization.cuh:72-87): cells with
// phi <= 0 (levelset::inside, levelset/levelset.hpp:166-170) take
// conservative(W).
void orc_set_states(const double* phi, const double* W, const orc_scheme* s, double* U,
                    int64_t cells, int nd) {
  const int S = 2 + nd;
  double c[4];
  orc_prim_to_cons(W, nd, 0, s, c);
  for (int64_t i = 0; i < cells; ++i)
    if (phi[i] <= 0.0)
      for (int k = 0; k < S; ++k) U[i * S + k] = c[k];
}

}  // extern "C"
