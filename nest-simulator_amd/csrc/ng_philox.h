// ng_philox.h -- the Philox_32 random stream of NEST's per-VP generators, for host and device.
//
// rng_type "Philox_32" (nestkernel/random_manager.cpp:69) is r123::Engine<r123::Philox4x32>
// (thirdparty/Random123/conventional/Engine.hpp over philox.h): a counter-based generator.
// The Philox4x32-10 block function (Salmon et al., "Parallel random numbers: as easy as 1, 2,
// 3", SC11) maps a 4-word counter and a 2-word key to 4 words; the engine increments the
// counter (starting from zero, so the first block has counter 1) and returns a block's words
// last word first. Word n of a stream is therefore a pure function of the key and n, which
// lets device threads draw any part of a stream independently.
#pragma once

#include <cstdint>

#if defined( __HIPCC__ )
#define NG_HD __host__ __device__
#else
#define NG_HD
#endif

namespace nestgpu
{

NG_HD inline void
philox4x32_10( const uint32_t ctr[ 4 ], uint32_t k0, uint32_t k1, uint32_t out[ 4 ] )
{
  uint32_t c0 = ctr[ 0 ], c1 = ctr[ 1 ], c2 = ctr[ 2 ], c3 = ctr[ 3 ];
  for ( int r = 0; r < 10; ++r )
  {
    if ( r > 0 )
    {
      k0 += 0x9E3779B9u;  // Weyl key schedule
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = static_cast< uint64_t >( 0xD2511F53u ) * c0;
    const uint64_t p1 = static_cast< uint64_t >( 0xCD9E8D57u ) * c2;
    const uint32_t n0 = static_cast< uint32_t >( p1 >> 32 ) ^ c1 ^ k0;
    const uint32_t n2 = static_cast< uint32_t >( p0 >> 32 ) ^ c3 ^ k1;
    c1 = static_cast< uint32_t >( p1 );
    c3 = static_cast< uint32_t >( p0 );
    c0 = n0;
    c2 = n2;
  }
  out[ 0 ] = c0;
  out[ 1 ] = c1;
  out[ 2 ] = c2;
  out[ 3 ] = c3;
}

// the block of engine words 4 b .. 4 b + 3 (counter b + 1)
NG_HD inline void
philox_block( uint32_t k0, uint32_t k1, uint64_t b, uint32_t out[ 4 ] )
{
  const uint64_t c = b + 1;
  const uint32_t ctr[ 4 ] = { static_cast< uint32_t >( c ), static_cast< uint32_t >( c >> 32 ), 0u, 0u };
  philox4x32_10( ctr, k0, k1, out );
}

// word n (0-based) of the stream
NG_HD inline uint32_t
philox_word( uint32_t k0, uint32_t k1, uint64_t n )
{
  uint32_t v[ 4 ];
  philox_block( k0, k1, n >> 2, v );
  const uint32_t i = 3u - static_cast< uint32_t >( n & 3u );
  return i == 0 ? v[ 0 ] : i == 1 ? v[ 1 ] : i == 2 ? v[ 2 ] : v[ 3 ];
}

// std::generate_canonical<double, 53> over a 32-bit engine (libstdc++ bits/random.tcc): two
// words, the first the low one, summed in double and scaled by 2^-64; a sum that rounds to
// 2^64 gives the largest double below 1
NG_HD inline double
philox_canonical( uint32_t lo, uint32_t hi )
{
  double s = static_cast< double >( lo );
  s += static_cast< double >( hi ) * 4294967296.0;
  const double u = s * 0x1p-64;
  return u < 1.0 ? u : 0x1.fffffffffffffp-1;
}

// Lemire's nearly divisionless draw in [0, range) as libstdc++'s uniform_int_distribution
// makes it from a 32-bit engine (bits/uniform_int_dist.h, _S_nd over 64-bit products): word w
// is rejected when the low half of w * range is below 2^32 mod range; an accepted word gives
// the high half. Acceptance depends on the word alone.
NG_HD inline bool
lemire_accepts( uint32_t w, uint32_t range )
{
  const uint32_t low = static_cast< uint32_t >( static_cast< uint64_t >( w ) * range );
  if ( low >= range )
    return true;
  const uint32_t threshold = ( 0u - range ) % range;
  return low >= threshold;
}

NG_HD inline uint32_t
lemire_value( uint32_t w, uint32_t range )
{
  return static_cast< uint32_t >( ( static_cast< uint64_t >( w ) * range ) >> 32 );
}

}  // namespace nestgpu
