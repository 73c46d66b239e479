// FlatParams.hpp -- internal to the drop-in library: a SimulationParameters flattened into the
// C ABI's mmp_params (which points into the vectors held here), and the status check that
// turns a C ABI error into MmpropError.
#ifndef MMPROP_DROPIN_FLAT_PARAMS_HPP
#define MMPROP_DROPIN_FLAT_PARAMS_HPP

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/mmprop.h"
#include "MultimodePropagation_GPU.hpp"

namespace mmprop_detail {

inline void check(mmp_status st, const char* where) {
  if (st != MMP_OK)
    throw MmpropError((int)st, std::string(where) + ": " + mmp_status_string(st) + " (" + mmp_last_error() + ")");
}

// Flat views of a SimulationParameters (the C ABI's mmp_params keeps pointers into these).
struct FlatParams {
  std::vector<double> beta, values, fields, response;
  std::vector<uint32_t> offsets, klm;
  mmp_params c{};

  FlatParams(const FlatParams&) = delete;  // c points into this object's vectors
  FlatParams& operator=(const FlatParams&) = delete;
  FlatParams(const SimulationParameters& in, bool raman) {
    const unsigned int M = in.n_modes, nt = in.nt;
    if (in.dispersion_coefficients.size() != M || in.coupling_coefficients.size() != M ||
        in.initial_fields.getNumberOfArrays() != M || in.initial_fields.getArraySize() != nt)
      throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "SimulationParameters sizes do not match n_modes and nt");
    const unsigned int nb = M ? (unsigned int)in.dispersion_coefficients[0].size() : 0;
    for (unsigned int p = 0; p < M; p++) {
      if (in.dispersion_coefficients[p].size() != nb)
        throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "every mode needs the same number of dispersion orders");
      beta.insert(beta.end(), in.dispersion_coefficients[p].begin(), in.dispersion_coefficients[p].end());
    }
    offsets.push_back(0);
    for (unsigned int p = 0; p < M; p++) {
      const Sparse3DArray& q = in.coupling_coefficients[p];
      for (unsigned int e = 0; e < q.getNumberOfNonZeroValues(); e++) {
        const NonZeroValue v = q.getNonZeroValue(e);
        klm.insert(klm.end(), {v.index[0], v.index[1], v.index[2]});
        values.push_back(v.value);
      }
      offsets.push_back((uint32_t)values.size());
    }
    fields.resize(2 * (size_t)M * nt);
    for (unsigned int p = 0; p < M; p++)
      for (unsigned int j = 0; j < nt; j++) {
        const std::complex<double> z = in.initial_fields.getValue(p, j);
        fields[2 * ((size_t)p * nt + j)] = z.real();
        fields[2 * ((size_t)p * nt + j) + 1] = z.imag();
      }
    if (raman) {
      if (in.raman_response.size() != nt)
        throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "raman_response must have nt samples");
      response = in.raman_response;
    }
    c.n_modes = M;
    c.fiber_length = in.fiber_length;
    c.n_beta = nb;
    c.dispersion_coefficients = beta.data();
    c.coupling_offsets = offsets.data();
    c.coupling_klm = klm.empty() ? nullptr : klm.data();
    c.coupling_values = values.empty() ? nullptr : values.data();
    c.nt = nt;
    c.time_window = in.time_window;
    c.pulse_width = in.pulse_width;
    c.initial_fields = fields.data();
    c.n_steps = in.n_steps;
    c.method_order = in.method_order;
    c.nonlinearity_const = in.nonlinearity_const;
    c.raman_proportion = in.raman_proportion;
    c.raman_response = raman ? response.data() : nullptr;
    c.raman = raman ? 1u : 0u;
  }
};


}  // namespace mmprop_detail

#endif
