// ng_aeif.h -- per-neuron adaptive Runge-Kutta-Fehlberg 4(5) for aeif_cond_exp.
//
// Stands in for the GSL stepper NEST uses (models/aeif_cond_exp.cpp:386,395,463:
// gsl_odeiv_step_rkf45, gsl_odeiv_control_yp_new(tol, tol), gsl_odeiv_evolve_apply).
// GSL is a third-party library outside the reference tree; its published method is:
//   * Fehlberg's 4(5) tableau, propagating the 5th-order solution, error = 5th - 4th;
//   * y' error control: D_i = eps_abs + eps_rel * |h y'_i|, r = max |yerr_i| / D_i,
//     shrink by S r^(-1/q) (>= 1/5) when r > 1.1, grow by S r^(-1/(q+1)) (<= 5, >= 1)
//     when r < 0.5, S = 0.9, q = 5 (the stepper's order);
//   * evolve: one accepted step towards t1 per call, clipped to t1, retried from the saved
//     state when the controller shrinks h and the shrink moves t by at least one ulp.
// The right-hand side follows aeif_cond_exp_dynamics (models/aeif_cond_exp.cpp:79-124).
#pragma once

#include "ng_math.h"

namespace nestgpu
{

struct AeifPar
{
  double V_peak, V_reset, g_L, C_m, E_ex, E_in, E_L, Delta_T, tau_w, a, b, V_th, tau_ex, tau_in, I_e, tol;
};

__host__ __device__ inline void
aeif_dynamics( const AeifPar& P, bool refractory, const double* y, double* f )
{
  const double V = refractory ? P.V_reset : ( P.V_peak < y[ 0 ] ? P.V_peak : y[ 0 ] );  // std::min
  const double g_ex = y[ 1 ];
  const double g_in = y[ 2 ];
  const double w = y[ 3 ];
  const double I_syn_exc = g_ex * ( V - P.E_ex );
  const double I_syn_inh = g_in * ( V - P.E_in );
  const double I_spike = P.Delta_T == 0. ? 0. : ( P.g_L * P.Delta_T * ng_exp( ( V - P.V_th ) / P.Delta_T ) );
  // currents_ carries no CurrentEvents on this path: I_stim_ == 0.0
  f[ 0 ] = refractory ? 0. : ( -P.g_L * ( V - P.E_L ) + I_spike - I_syn_exc - I_syn_inh - w + P.I_e + 0.0 ) / P.C_m;
  f[ 1 ] = -g_ex / P.tau_ex;
  f[ 2 ] = -g_in / P.tau_in;
  f[ 3 ] = ( P.a * ( V - P.E_L ) - w ) / P.tau_w;
}

__host__ __device__ inline void
aeif_rkf45_step( const AeifPar& P, bool refr, double* y, double h, const double* k1, double* yerr, double* dydt_out )
{
  const double b21 = 1.0 / 4.0;
  const double b31 = 3.0 / 32.0, b32 = 9.0 / 32.0;
  const double b41 = 1932.0 / 2197.0, b42 = -7200.0 / 2197.0, b43 = 7296.0 / 2197.0;
  const double b51 = 8341.0 / 4104.0, b52 = -32832.0 / 4104.0, b53 = 29440.0 / 4104.0, b54 = -845.0 / 4104.0;
  const double b61 = -6080.0 / 20520.0, b62 = 41040.0 / 20520.0, b63 = -28352.0 / 20520.0, b64 = 9295.0 / 20520.0,
               b65 = -5643.0 / 20520.0;
  const double c1 = 902880.0 / 7618050.0, c3 = 3953664.0 / 7618050.0, c4 = 3855735.0 / 7618050.0,
               c5 = -1371249.0 / 7618050.0, c6 = 277020.0 / 7618050.0;
  const double ec1 = 1.0 / 360.0, ec3 = -128.0 / 4275.0, ec4 = -2197.0 / 75240.0, ec5 = 1.0 / 50.0, ec6 = 2.0 / 55.0;
  double k2[ 4 ], k3[ 4 ], k4[ 4 ], k5[ 4 ], k6[ 4 ], yt[ 4 ];
#pragma unroll
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + b21 * h * k1[ i ];
  aeif_dynamics( P, refr, yt, k2 );
#pragma unroll
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b31 * k1[ i ] + b32 * k2[ i ] );
  aeif_dynamics( P, refr, yt, k3 );
#pragma unroll
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b41 * k1[ i ] + b42 * k2[ i ] + b43 * k3[ i ] );
  aeif_dynamics( P, refr, yt, k4 );
#pragma unroll
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b51 * k1[ i ] + b52 * k2[ i ] + b53 * k3[ i ] + b54 * k4[ i ] );
  aeif_dynamics( P, refr, yt, k5 );
#pragma unroll
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b61 * k1[ i ] + b62 * k2[ i ] + b63 * k3[ i ] + b64 * k4[ i ] + b65 * k5[ i ] );
  aeif_dynamics( P, refr, yt, k6 );
#pragma unroll
  for ( int i = 0; i < 4; ++i )
  {
    const double di = c1 * k1[ i ] + c3 * k3[ i ] + c4 * k4[ i ] + c5 * k5[ i ] + c6 * k6[ i ];
    y[ i ] = y[ i ] + h * di;
  }
#pragma unroll
  for ( int i = 0; i < 4; ++i )
  {
    const double ei = ec1 * k1[ i ] + ec3 * k3[ i ] + ec4 * k4[ i ] + ec5 * k5[ i ] + ec6 * k6[ i ];
    yerr[ i ] = h * ei;
  }
  aeif_dynamics( P, refr, y, dydt_out );
}

// returns -1 (decrease), 0 (keep), +1 (increase)
__host__ __device__ inline int
aeif_yp_hadjust( double tol, const double* yerr, const double* yp, double* h )
{
  const double S = 0.9;
  const double h_old = *h;
  double rmax = 2.2250738585072014e-308;  // DBL_MIN
#pragma unroll
  for ( int i = 0; i < 4; ++i )
  {
    const double D0 = tol * ( 0.0 * fabs( 0.0 ) + 1.0 * fabs( h_old * yp[ i ] ) ) + tol;
    const double r = fabs( yerr[ i ] ) / fabs( D0 );
    rmax = r > rmax ? r : rmax;
  }
  if ( rmax > 1.1 )
  {
    double r = S / ng_pow( rmax, 1.0 / 5.0 );
    if ( r < 0.2 )
      r = 0.2;
    *h = r * h_old;
    return -1;
  }
  else if ( rmax < 0.5 )
  {
    double r = S / ng_pow( rmax, 1.0 / ( 5.0 + 1.0 ) );
    if ( r > 5.0 )
      r = 5.0;
    if ( r < 1.0 )
      r = 1.0;
    *h = r * h_old;
    return 1;
  }
  return 0;
}

// One accepted adaptive step from *t towards t1 (gsl_odeiv_evolve_apply semantics).
__host__ __device__ inline void
aeif_evolve( const AeifPar& P, bool refr, double* t, double t1, double* h, double* y )
{
  const double t0 = *t;
  double h0 = *h;
  const double dt = t1 - t0;
  double y0[ 4 ], dydt_in[ 4 ], yerr[ 4 ], dydt_out[ 4 ];
#pragma unroll
  for ( int i = 0; i < 4; ++i )
    y0[ i ] = y[ i ];
  aeif_dynamics( P, refr, y, dydt_in );
  // A rejected attempt shrinks h0 by a factor of at most 0.9 * 1.1^(-1/5) < 0.89 and stops
  // once t + h0 rounds to t, so fewer than 400 attempts can occur; the bound only makes the
  // loop's exit evident.
  for ( int attempt = 0; attempt < 10000; ++attempt )
  {
    bool final_step = false;
    if ( h0 > dt )
    {
      h0 = dt;
      final_step = true;
    }
    aeif_rkf45_step( P, refr, y, h0, dydt_in, yerr, dydt_out );
    *t = final_step ? t1 : t0 + h0;
    const double h_old = h0;
    const int st = aeif_yp_hadjust( P.tol, yerr, dydt_out, &h0 );
    if ( st < 0 )
    {
      const double t_curr = *t;
      const double t_next = ( *t ) + h0;
      if ( fabs( h0 ) < fabs( h_old ) and t_next != t_curr )
      {
#pragma unroll
        for ( int i = 0; i < 4; ++i )
          y[ i ] = y0[ i ];
        continue;
      }
      h0 = h_old;
    }
    break;
  }
  *h = h0;
}

}  // namespace nestgpu
