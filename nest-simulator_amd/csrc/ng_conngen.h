// ng_conngen.h -- connectivity generated on the device from counter-based per-VP streams.
//
// NEST's builders draw every random number of a Connect call from the stream of the VP that
// owns the target (FixedInDegreeBuilder::inner_connect_ conn_builder.cpp:1508-1545,
// BernoulliBuilder::inner_connect_ :1937-1965). With rng_type "Philox_32" word n of a stream is
// a function of its key and n alone, so when the number of words each draw takes is known in
// advance, device threads can make all draws of a call at once and still produce the
// connections, in creation order, that the serial builder makes:
//   - pairwise_bernoulli: one drand (two words) per (target, source) pair the builder visits
//     (an autapse it skips takes none), and no parameter draws (constant weight and delay);
//   - fixed_indegree with autapses and multapses allowed: one uniform_int draw per
//     connection, each taking the next word that libstdc++'s Lemire method accepts; whether
//     a word is accepted depends on the word alone, so the rejected positions of the stream
//     are found first and every draw then knows its word.
#pragma once

#include <cstdint>
#include <vector>

namespace nestgpu
{

// a per-VP stream: Philox key and the position of its next word
struct GenStream
{
  uint32_t k0, k1;
  uint64_t pos;
};

// one local target of a VP, in the VP's builder order
struct GenRow
{
  uint32_t tgt;    // target node id
  int32_t aut;     // pairwise_bernoulli: index in the sources of the autapse skipped, or -1
  uint64_t draw0;  // pairwise_bernoulli: the row's first draw in its VP's call
  uint32_t vp;
};

// pairwise_bernoulli over pre (node ids) for the rows; src[vp] / tgt[vp] receive each VP's
// connections in creation order. The streams are not advanced (the caller knows the draws).
void gen_pairwise_bernoulli( int device, const std::vector< uint32_t >& pre, const std::vector< GenRow >& rows,
  const std::vector< GenStream >& streams, double p, std::vector< std::vector< uint32_t > >& src,
  std::vector< std::vector< uint32_t > >& tgt );

// fixed_indegree (autapses and multapses allowed): row_tgt[vp] are the local targets of each
// VP in builder order, K draws each. src[vp] / tgt[vp] receive the connections; each stream's
// position is advanced past the words the draws took.
void gen_fixed_indegree( int device, const std::vector< uint32_t >& pre,
  const std::vector< std::vector< uint32_t > >& row_tgt, uint32_t K, std::vector< GenStream >& streams,
  std::vector< std::vector< uint32_t > >& src, std::vector< std::vector< uint32_t > >& tgt );

}  // namespace nestgpu
