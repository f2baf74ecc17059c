// Checker probe (test infrastructure): seeds std::mt19937_64 through the reference's own
// seed sequence header exactly as nestkernel/random_generators.h:155-160 does, and prints
// the seed-sequence words and the first engine outputs for the seeds given on the command
// line. Built from the reference tree by oracle/ref/Makefile into oracle/_ref/.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "randutils.hpp"

int
main( int argc, char** argv )
{
  // usage: randutils_probe n_words n_draws seed0 [seed1 ...]
  if ( argc < 4 )
    return 2;
  const int n_words = std::atoi( argv[ 1 ] );
  const int n_draws = std::atoi( argv[ 2 ] );
  std::vector< std::uint32_t > s;
  for ( int i = 3; i < argc; ++i )
    s.push_back( static_cast< std::uint32_t >( std::strtoul( argv[ i ], nullptr, 0 ) ) );
  std::vector< std::uint32_t > words( n_words );
  {
    randutils::seed_seq_fe128 sseq( s.begin(), s.end() );
    sseq.generate( words.begin(), words.end() );
  }
  for ( auto w : words )
    std::printf( "%u\n", w );
  randutils::seed_seq_fe128 sseq( s.begin(), s.end() );
  std::mt19937_64 eng;
  eng.seed( sseq );
  for ( int i = 0; i < n_draws; ++i )
    std::printf( "%llu\n", static_cast< unsigned long long >( eng() ) );
  return 0;
}
