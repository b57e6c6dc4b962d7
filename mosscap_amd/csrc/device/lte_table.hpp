// Tabulated LTE hydrogen EOS (eos.type = tabulatedlteh): the update_eos pass
// of TabulatedLteH.hpp:57-161 as a device kernel. Before every stage each
// cell (ghosts included) samples the table's ionisation fraction y and
// log10 T at (log10 eint/rho, log10 rho) by clamped bilinear interpolation,
// keeps them as fields, and adds the change of ionisation energy since the
// previous stage, rho chi_H / m_H (y - y_prev), to E (nothing on the first
// call, where y_prev = -1).
#pragma once

#include "functors.hpp"
#include "lte_eos.hpp"

namespace mh {

struct TabParams {
    double ex_start, ex_step;  // log10(eint / rho) axis
    int32_t ex_num;
    double r_start, r_step;    // log10(rho) axis
    int32_t r_num;
};

// TabulatedLteH::sample, TabulatedLteH.hpp:57-118. Tables are [r_num][ex_num].
// The floor-based index is formed only inside the table, so a NaN coordinate
// takes the first edge and every access stays in the table; other inputs give
// the reference's indices.
__device__ inline void tab_sample(const TabParams& t, const double* __restrict__ ytab,
                                  const double* __restrict__ ltab, double log_eint_rho, double log_rho, double& y,
                                  double& log_T) {
    const double frac_x = (log_eint_rho - t.ex_start) / t.ex_step;
    int ix;
    int ixp;
    double tx, txp;
    const int max_x = t.ex_num - 1;
    const int max_y = t.r_num - 1;
    if (!(frac_x >= 0.0) || frac_x >= max_x) {  // outside the table (or NaN): its edge
        ix = frac_x >= max_x ? max_x : 0;
        ixp = ix;
        tx = 1.0;
        txp = 0.0;
    } else {
        ix = int(floor(frac_x));
        ixp = ix + 1;
        txp = frac_x - ix;
        tx = 1.0 - txp;
    }
    const double frac_y = (log_rho - t.r_start) / t.r_step;
    int iy;
    int iyp;
    double ty, typ;
    if (!(frac_y >= 0.0) || frac_y >= max_y) {
        iy = frac_y >= max_y ? max_y : 0;
        iyp = iy;
        ty = 1.0;
        typ = 0.0;
    } else {
        iy = int(floor(frac_y));
        iyp = iy + 1;
        typ = frac_y - iy;
        ty = 1.0 - typ;
    }
    const int r0 = iy * t.ex_num, r1 = iyp * t.ex_num;
    y = ty * (tx * ytab[r0 + ix] + txp * ytab[r0 + ixp]) + typ * (tx * ytab[r1 + ix] + txp * ytab[r1 + ixp]);
    log_T = ty * (tx * ltab[r0 + ix] + txp * ltab[r0 + ixp]) + typ * (tx * ltab[r1 + ix] + txp * ltab[r1 + ixp]);
}

// update_eos over every cell of `in`, result into `out` (the same buffer, or
// a copy of all nvar planes with E updated). A cell with eint < 0 sets
// *eint_negative (the reference aborts) and is left unchanged.
template <class Form, int ND>
__global__ __launch_bounds__(256) void tab_eos_kernel(typename Form::Planes in, typename Form::Planes out, int ncells,
                                                      int nvar, TabParams t, const double* __restrict__ ytab,
                                                      const double* __restrict__ ltab, double* y_space,
                                                      double* T_space, int* eint_negative) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    constexpr int IE = 1 + ND;
    const bool copy = Form::get_ptr(in) != Form::get_ptr(out);
    if (copy) {
        for (int v = 0; v < nvar; ++v)
            if (v != IE) Form::put(out, v * ncells + c, Form::get(in, v * ncells + c));
    }
    const auto e_raw = Form::get(in, IE * ncells + c);
    double mom2_sum = square(Form::value(Form::get(in, ncells + c)));
    if constexpr (ND > 1) mom2_sum += square(Form::value(Form::get(in, 2 * ncells + c)));
    if constexpr (ND > 2) mom2_sum += square(Form::value(Form::get(in, 3 * ncells + c)));
    const double rho = Form::value(Form::get(in, c));
    const double e_kin = 0.5 * mom2_sum / rho;
    const double eint = Form::value(e_raw) - e_kin;
    if (eint < 0.0) {
        atomicOr(eint_negative, 1);
        if (copy) Form::put(out, IE * ncells + c, e_raw);
        return;
    }
    const double log_rho = log10(rho);
    const double log_eint_rho = log10(eint) - log_rho;
    double y, log_T;
    tab_sample(t, ytab, ltab, log_eint_rho, log_rho, y, log_T);
    const double prev_y = y_space[c];
    y_space[c] = y;
    const double delta_eint = (prev_y < 0.0) ? 0.0 : rho * (lte::kChiH / lte::kHMass) * (y - prev_y);
    T_space[c] = pow(10.0, log_T);
    Form::put(out, IE * ncells + c, Form::add_double(e_raw, delta_eint));
}

}  // namespace mh
