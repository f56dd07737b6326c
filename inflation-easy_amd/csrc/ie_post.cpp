// ie_post.cpp -- curvature and tensor power spectra from the binned spectra.
//
// The reference computes these in its analysis notebook from the written files
// (notebooks/plot.ipynb, cells 9 and 23): each spectrum row pairs with the
// modes.dat row of the same index, k = that row's value.
//   P_zeta^lin(k)   = P_phi(k) k^3 / (2 pi^2) (phidot / H)^-2   (cell 9)
//   P_zeta^dN(k)    = P_dN(k) k^3 / (2 pi^2)                      (cell 9)
//   Delta_h^2(k)    = P_h(k) k^3 / (2 pi^2)                       (cell 23)
//   Omega_GW(k)     = k^3 / (2 pi^2) (P_hdot + (k/a)^2 P_h) / (12 H^2)  (cell 23)
// with phidot the velocity.dat value and H, a the sf.dat values at that output.
#include <math.h>

#include "ie_api.h"

namespace {

const double kTwoPi2 = 2.0 * M_PI * M_PI;

bool bad(int nbins, const double *k, const void *p, const double *out) { return nbins < 0 || !k || !p || !out; }

}  // namespace

extern "C" {

ie_status ie_pzeta_linear(int nbins, const double *k, const double *p_phi, double phidot, double hubble,
                          double *out)
{
    if (bad(nbins, k, p_phi, out)) return IE_ERR_ARG;
    const double r = pow(phidot / hubble, -2.0);
    for (int b = 0; b < nbins; ++b) out[b] = p_phi[b] * (k[b] * k[b] * k[b]) / kTwoPi2 * r;
    return IE_OK;
}

ie_status ie_pzeta_deltaN(int nbins, const double *k, const double *p_dn, double *out)
{
    if (bad(nbins, k, p_dn, out)) return IE_ERR_ARG;
    for (int b = 0; b < nbins; ++b) out[b] = p_dn[b] * (k[b] * k[b] * k[b]) / kTwoPi2;
    return IE_OK;
}

ie_status ie_delta_h2(int nbins, const double *k, const double *p_h, double *out)
{
    if (bad(nbins, k, p_h, out)) return IE_ERR_ARG;
    for (int b = 0; b < nbins; ++b) out[b] = (k[b] * k[b] * k[b]) / kTwoPi2 * p_h[b];
    return IE_OK;
}

ie_status ie_omega_gw(int nbins, const double *k, const double *p_h, const double *p_hdot, double a, double hubble,
                      double *out)
{
    if (bad(nbins, k, p_h, out) || !p_hdot) return IE_ERR_ARG;
    for (int b = 0; b < nbins; ++b) {
        const double term2 = (k[b] / a) * (k[b] / a) * p_h[b];
        out[b] = (1.0 / (12.0 * hubble * hubble)) * ((k[b] * k[b] * k[b]) / kTwoPi2) * (p_hdot[b] + term2);
    }
    return IE_OK;
}

}  // extern "C"
