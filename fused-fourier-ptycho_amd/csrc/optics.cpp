// optics.cpp -- LED illumination geometry -> integer spectrum windows, FP64.
//
// Restates codes_v2/func_ware/init_environment_rgb.m:33-134 and the index
// assembly of codes_v2/main_run_FPM.m:23,27-31,53-55:
//   Fxx1(k) = ((k - 1 - NN/2) / (n * dpix)) * 2 pi,   k = 1..NN, NN = n * pratio
//   LED (i, j) of a g x g grid, centre c = (g - 1)/2, v_i = i - 1 - c, h_j = j - 1 - c:
//     den = sqrt((-v_i D)^2 + (-h_j D)^2 + h^2)
//     u = -v_i D / den,  v = h_j D / den      (the swap at init_environment_rgb.m:86-88)
//   uo = argmin_k |Fxx1(k) - k0 u|,  vo = argmin_k |Fyy1(k) - k0 v|  (1-based)
//   kl = -n/2 + uo,  kt = -n/2 + vo
// LEDs are listed in the centre-out spiral of init_environment_rgb.m:110-134
// (or in grid order).  The argmin is required to be unique by a relative
// margin of 1e-9, so the integers do not depend on the last bits of the FP64
// arithmetic (MATLAB's find(min) would return several indices on a tie).
#include "../../include/fpm_amd.h"

#include <cmath>
#include <string>
#include <vector>

namespace fpm {
void set_error(const std::string& msg);
}

namespace {

// 1-based argmin of |grid(k) - target| over k = 1..NN; -1 on a near tie
int argmin_unique(const std::vector<double>& grid, double target) {
    int best = -1;
    double bv = INFINITY, second = INFINITY;
    for (size_t k = 0; k < grid.size(); ++k) {
        const double d = std::fabs(grid[k] - target);
        if (d < bv) {
            second = bv;
            bv = d;
            best = (int)k + 1;
        } else if (d < second) {
            second = d;
        }
    }
    const double scale = std::fabs(grid[1] - grid[0]);
    if (!(second - bv > 1e-9 * scale)) return -1;
    return best;
}

}  // namespace

// LED (i, j) seen from the field area at (objdx, objdy) of the sample plane:
// the lateral offsets above become (-v_i D + objdx, -h_j D + objdy)
// (init_environment_rgb.m:26-29,83-84, where the reference fixes objdx =
// objdy = 0).
static int led_offsets_from(double objdx_um, double objdy_um, double wavelength_um, double na, double pixel_um,
                            double led_pitch_um, double led_height_um, int32_t led_grid, int64_t n, int32_t pratio,
                            int32_t spiral, int32_t* led_idx, int32_t* grid_index, double* na_illum) {
    fpm::set_error("");
    (void)na;  // the pupil radius does not enter the window offsets
    if (!led_idx) {
        fpm::set_error("fpm_led_offsets: null led_idx");
        return FPM_E_INVALID;
    }
    if (led_grid < 1 || (led_grid % 2) == 0 || n < 2 || (n % 2) != 0 || pratio < 1 || !(wavelength_um > 0) ||
        !(pixel_um > 0) || !(led_height_um > 0)) {
        fpm::set_error("fpm_led_offsets: invalid optics (odd led_grid >= 1, even n, pratio >= 1, positive lengths)");
        return FPM_E_INVALID;
    }
    const double two_pi = 2.0 * 3.14159265358979323846;
    const double k0 = two_pi / wavelength_um;
    const int64_t NN = n * (int64_t)pratio;
    std::vector<double> F(NN);
    for (int64_t k = 0; k < NN; ++k) F[k] = ((double)(k - NN / 2) / ((double)n * pixel_um)) * two_pi;

    const int g = led_grid;
    const double c = (g - 1) / 2.0;
    std::vector<int> uo(g * g), vo(g * g);
    std::vector<double> nai(g * g);
    for (int i = 0; i < g; ++i) {
        for (int j = 0; j < g; ++j) {
            const double vv = (double)i - c, hh = (double)j - c;
            const double a = -vv * led_pitch_um + objdx_um, b = -hh * led_pitch_um + objdy_um;
            const double den = std::sqrt(a * a + b * b + led_height_um * led_height_um);
            const double uu = a / den;    // -v_i D / den
            const double v = -b / den;    //  h_j D / den
            const int x = argmin_unique(F, k0 * uu), y = argmin_unique(F, k0 * v);
            if (x < 0 || y < 0) {
                fpm::set_error("fpm_led_offsets: LED (" + std::to_string(i + 1) + "," + std::to_string(j + 1) +
                               ") has no unique nearest spectrum index (tie)");
                return FPM_E_INVALID;
            }
            uo[i * g + j] = x;
            vo[i * g + j] = y;
            nai[i * g + j] = std::sqrt(uu * uu + v * v);
        }
    }

    // visiting order: (row, col) 0-based
    std::vector<int> oi, oj;
    oi.reserve(g * g);
    oj.reserve(g * g);
    if (spiral) {
        // init_environment_rgb.m:110-134 with sin/cos(pi/2 d) taken exactly
        static const int S[4] = {0, 1, 0, -1}, C[4] = {1, 0, -1, 0};
        std::vector<char> seen(g * g, 0);
        int ii = (g + 1) / 2 - 1, jj = (g + 1) / 2 - 1, dir = 0;
        oi.push_back(ii);
        oj.push_back(jj);
        seen[ii * g + jj] = 1;
        for (int cnt = 1; cnt < g * g; ++cnt) {
            const int d2 = dir + 1;
            const int i2 = ii + S[d2 & 3], j2 = jj + C[d2 & 3];
            if (i2 >= 0 && i2 < g && j2 >= 0 && j2 < g && !seen[i2 * g + j2]) dir = d2;
            ii += S[dir & 3];
            jj += C[dir & 3];
            if (ii < 0 || ii >= g || jj < 0 || jj >= g || seen[ii * g + jj]) {
                fpm::set_error("fpm_led_offsets: spiral left the LED grid");
                return FPM_E_INVALID;
            }
            oi.push_back(ii);
            oj.push_back(jj);
            seen[ii * g + jj] = 1;
        }
    } else {
        for (int i = 0; i < g; ++i)
            for (int j = 0; j < g; ++j) {
                oi.push_back(i);
                oj.push_back(j);
            }
    }
    for (int l = 0; l < g * g; ++l) {
        const int q = oi[l] * g + oj[l];
        led_idx[2 * l + 0] = (int32_t)(-n / 2 + uo[q]);  // kl
        led_idx[2 * l + 1] = (int32_t)(-n / 2 + vo[q]);  // kt
        if (grid_index) grid_index[l] = (oj[l] + 1) + oi[l] * g;  // k(i,j) = j + (i-1) g
        if (na_illum) na_illum[l] = nai[q];
    }
    return FPM_OK;
}

extern "C" int fpm_led_offsets(double wavelength_um, double na, double pixel_um, double led_pitch_um,
                               double led_height_um, int32_t led_grid, int64_t n, int32_t pratio, int32_t spiral,
                               int32_t* led_idx, int32_t* grid_index, double* na_illum) {
    return led_offsets_from(0.0, 0.0, wavelength_um, na, pixel_um, led_pitch_um, led_height_um, led_grid, n, pratio,
                            spiral, led_idx, grid_index, na_illum);
}

extern "C" int fpm_led_offsets_at(double objdx_um, double objdy_um, double wavelength_um, double na,
                                  double pixel_um, double led_pitch_um, double led_height_um, int32_t led_grid,
                                  int64_t n, int32_t pratio, int32_t spiral, int32_t* led_idx, int32_t* grid_index,
                                  double* na_illum) {
    if (!std::isfinite(objdx_um) || !std::isfinite(objdy_um)) {
        fpm::set_error("fpm_led_offsets_at: non-finite field offset");
        return FPM_E_INVALID;
    }
    return led_offsets_from(objdx_um, objdy_um, wavelength_um, na, pixel_um, led_pitch_um, led_height_um, led_grid, n,
                            pratio, spiral, led_idx, grid_index, na_illum);
}
