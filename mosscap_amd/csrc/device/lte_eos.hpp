// Analytic LTE hydrogen EOS fields (eos.type = analyticlteh): the update_eos
// pass of AnalyticLteH.hpp:51-125 as a device kernel. Per cell, the
// temperature and ionisation fraction of a pure-hydrogen gas in LTE with the
// cell's density and internal energy, by bisection on T with the Saha
// equation for y. The hydro step does not read them (they are output fields);
// the arithmetic and the bisection's stopping rules follow the reference.
#pragma once

#include "functors.hpp"

namespace mh {

struct LteParams {
    double gamma;
    double inv_avg_mass;
    int32_t ionisation_e;  // include the ionisation energy in the internal energy
};

namespace lte {
constexpr double kHMass = 1.6737830080950003e-27;  // [kg]
constexpr double kKB = 1.380649e-23;               // [J/K]
constexpr double kChiH = 2.178710282685096e-18;    // [J]

__device__ inline double saha_rhs_h(double temp) {
    return 2.4146830395719654e+21 * pow(temp, 1.5) * exp(-157763.42386247337 / temp);
}
__device__ inline double y_from_nhtot(double nhtot, double temp) {
    const double X = saha_rhs_h(temp);
    return 0.5 * (-X + sqrt(X * X + 4 * nhtot * X)) / nhtot;
}
}  // namespace lte

template <class Form, int ND>
__global__ __launch_bounds__(256) void lte_eos_kernel(typename Form::Planes q, int ncells, LteParams prm,
                                                      double* __restrict__ y_out, double* __restrict__ T_out,
                                                      int* flipped) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const double rho = Form::value(Form::get(q, c));
    double mom2_sum = square(Form::value(Form::get(q, ncells + c)));
    if constexpr (ND > 1) mom2_sum += square(Form::value(Form::get(q, 2 * ncells + c)));
    if constexpr (ND > 2) mom2_sum += square(Form::value(Form::get(q, 3 * ncells + c)));
    const double e_kin = 0.5 * mom2_sum / rho;
    const double eint = Form::value(Form::get(q, (1 + ND) * ncells + c)) - e_kin;
    const double e_to_T = (prm.gamma - 1.0) / (rho * (lte::kKB / lte::kHMass) * prm.inv_avg_mass);
    const double ion = prm.ionisation_e ? 1.0 : 0.0;
    auto temp_from_y = [&](double yy) {
        return e_to_T / (1.0 + yy) * (eint - ion * yy * rho * (lte::kChiH / lte::kHMass) * prm.inv_avg_mass);
    };
    const double rho_to_nhtot = (1.0 / lte::kHMass) * prm.inv_avg_mass;
    const double t1 = temp_from_y(1.0);
    const double tb0 = t1 > 100.0 ? t1 : 100.0, tb1 = temp_from_y(0.0);
    if (tb0 > tb1) {
        atomicOr(flipped, 1);
        y_out[c] = T_out[c] = __builtin_nan("");
        return;
    }
    double temp_step = tb1 - tb0, temp = tb0, y = 0.0;
    for (int iter = 0; iter < 100; ++iter) {
        temp_step *= 0.5;
        const double test_temp = temp + temp_step;
        y = lte::y_from_nhtot(rho * rho_to_nhtot, test_temp);
        const double temp_err = test_temp - temp_from_y(y);
        if (temp_err <= 0.0) temp = test_temp;
        if (__builtin_fabs(temp_step) < 1e-1 || __builtin_fabs(temp_err) < 1e-1) break;
    }
    y_out[c] = y;
    T_out[c] = temp;
}

}  // namespace mh
