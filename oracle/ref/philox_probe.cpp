// Checker probe (test infrastructure): seeds r123::Engine<r123::Philox4x32> through the
// reference's own Random123 and seed-sequence headers exactly as nestkernel/
// random_generators.h:155-160 does for rng_type "Philox_32" (random_manager.cpp:69), and
// prints draws of one family for the seed words given on the command line. Built from the
// reference tree by oracle/ref/Makefile into oracle/_ref/.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "Random123/conventional/Engine.hpp"
#include "Random123/philox.h"
#include "randutils.hpp"

int
main( int argc, char** argv )
{
  // usage: philox_probe family n range seed0 [seed1 ...]
  //   family 0 engine words, 1 uniform_real(0, 1), 2 uniform_int(0, range - 1), 3 normal(0, 1)
  if ( argc < 5 )
    return 2;
  const int family = std::atoi( argv[ 1 ] );
  const int n = std::atoi( argv[ 2 ] );
  const unsigned long range = std::strtoul( argv[ 3 ], nullptr, 0 );
  std::vector< std::uint32_t > s;
  for ( int i = 4; i < argc; ++i )
    s.push_back( static_cast< std::uint32_t >( std::strtoul( argv[ i ], nullptr, 0 ) ) );
  randutils::seed_seq_fe128 sseq( s.begin(), s.end() );
  r123::Engine< r123::Philox4x32 > eng;
  eng.seed( sseq );
  std::uniform_real_distribution<> u01( 0.0, 1.0 );
  std::uniform_int_distribution< unsigned long > ul;
  std::normal_distribution<> nd( 0.0, 1.0 );
  for ( int i = 0; i < n; ++i )
  {
    if ( family == 0 )
      std::printf( "%u\n", static_cast< unsigned >( eng() ) );
    else if ( family == 1 )
      std::printf( "%.17g\n", u01( eng ) );
    else if ( family == 2 )
    {
      std::uniform_int_distribution< unsigned long >::param_type p( 0, range - 1 );
      std::printf( "%lu\n", ul( eng, p ) );
    }
    else
      std::printf( "%.17g\n", nd( eng ) );
  }
  return 0;
}
