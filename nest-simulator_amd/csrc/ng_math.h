// ng_math.h -- portable double exp/log for the adaptive aeif integrator.
//
// Method: Sun fdlibm e_exp.c / e_log.c (argument reduction by ln 2; rational kernel for
// exp, polynomial in s = f/(2+f) for log). Every operation is an IEEE add, multiply or
// divide in a fixed order, so host and device builds compiled without contraction return
// identical bits. Accuracy is within 1 ulp of the correctly rounded result.
#pragma once

#include <cmath>
#include <limits>

namespace nestgpu
{

__host__ __device__ inline double
ng_exp( double x )
{
  const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10,
               invln2 = 1.44269504088896338700e+00;
  const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
               P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
               P5 = 4.13813679705723846039e-08;
  if ( x != x )
    return x;
  if ( x > 709.782712893383973096 )
    return INFINITY;
  if ( x < -745.13321910194110842 )
    return 0.0;
  const double kd = floor( x * invln2 + 0.5 );
  const int k = static_cast< int >( kd );
  const double hi = x - kd * ln2hi;
  const double lo = kd * ln2lo;
  const double rr = hi - lo;
  const double t = rr * rr;
  const double c = rr - t * ( P1 + t * ( P2 + t * ( P3 + t * ( P4 + t * P5 ) ) ) );
  const double y = 1.0 - ( ( lo - ( rr * c ) / ( 2.0 - c ) ) - hi );
  return ldexp( y, k );
}

__host__ __device__ inline double
ng_log( double x )
{
  const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10;
  const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01, Lg3 = 2.857142874366239149e-01,
               Lg4 = 2.222219843214978396e-01, Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
               Lg7 = 1.479819860511658591e-01;
  if ( !( x > 0 ) )
    return x == 0 ? -INFINITY : NAN;
  if ( x == INFINITY )
    return x;
  int e;
  double mnt = frexp( x, &e );
  if ( mnt < 0.70710678118654752440 )
  {
    mnt = mnt * 2.0;
    e -= 1;
  }
  const double f = mnt - 1.0;
  const double kd = static_cast< double >( e );
  const double s = f / ( 2.0 + f );
  const double z = s * s;
  const double w = z * z;
  const double t1 = w * ( Lg2 + w * ( Lg4 + w * Lg6 ) );
  const double t2 = z * ( Lg1 + w * ( Lg3 + w * ( Lg5 + w * Lg7 ) ) );
  const double R = t2 + t1;
  const double hfsq = 0.5 * f * f;
  return kd * ln2hi - ( ( hfsq - ( s * ( hfsq + R ) + kd * ln2lo ) ) - f );
}

__host__ __device__ inline double
ng_pow( double x, double y )
{
  return ng_exp( y * ng_log( x ) );
}

}  // namespace nestgpu
