/*
 * chns_tables_ref.c - ETD2RK coefficients of one mode for the C restatement
 * (oracle/chns_step.c; TEST INFRASTRUCTURE).  Compiled without FP contraction
 * (-ffp-contract=off) so each expression rounds as the reference's
 * unoptimised build does: the phi_2 coefficient cancels catastrophically at
 * low k, and only the reference's operation order reproduces its values.
 * Follows etd2rk.c:88-127 (stage 1) and :162-203 (stage 2): thresholds
 * |c| > 1e-12 (binary stage 2: c != 0.0), fallbacks dt and 0.5*dt.
 */
#include <math.h>

void chnso_etd_mode(int model, double nu, double dt, double lambda, double epsilon, double sigma, double ksqr,
                    double *eta_o, double *z1w_o, double *z2w_o, double *beta_o, double *z1c_o, double *z2c_o) {
  double c = -nu * ksqr;                              /* etd2rk.c:90,164 */
  const double eta = exp(-nu * dt * ksqr);            /* :91,165 */
  *eta_o = eta;
  *z1w_o = fabs(c) > 1e-12 ? (eta - 1.0) / c : dt;    /* :92-97 */
  *z2w_o = fabs(c) > 1e-12 ? (eta - 1.0 - c * dt) / (dt * c * c) : 0.5 * dt;   /* :167-172 */
  *beta_o = *z1c_o = *z2c_o = 0.0;
  if (model == 0) return;
  int big2;
  if (model == 1) {
    c = -1.5 * lambda * sigma * epsilon * ksqr * ksqr;   /* :102,177 */
    big2 = fabs(c) != 0.0;                               /* :181 */
  } else {
    c = -0.75 * lambda * epsilon;                        /* :117,191 */
    c *= ksqr * ksqr;                                    /* :118,192 */
    big2 = fabs(c) > 1e-12;                              /* :195 */
  }
  const double beta = exp(dt * c);                       /* :104,119,179,193 */
  *beta_o = beta;
  *z1c_o = fabs(c) > 1e-12 ? (beta - 1.0) / c : dt;      /* :106-111,121-127 */
  if (big2) {
    double z = (beta - 1.0 - c * dt);                    /* :183,197 */
    z /= (dt * c * c);                                   /* :183,198 */
    *z2c_o = z;
  } else {
    *z2c_o = 0.5 * dt;
  }
}
