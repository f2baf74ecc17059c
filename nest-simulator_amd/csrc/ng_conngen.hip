// ng_conngen.hip -- gfx950 kernels of device-side connectivity generation (see ng_conngen.h).
//
// pairwise_bernoulli: the pairs of a call are cut into segments of one target and SEG
// consecutive sources; a first pass counts each segment's connections, a scan on the host
// gives every segment its output offset, and a second pass writes the connections. Each
// thread takes SPT consecutive sources, so its draws are consecutive words of the stream and
// one Philox block serves two draws.
//
// fixed_indegree: a first pass marks the rejected words of each VP's stream over the window
// the call can reach; with the sorted rejected positions r_0 < r_1 < ..., draw q takes word
// pos + q + m, m the number of i with r_i - pos - i <= q.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "ng_conngen.h"
#include "ng_engine.h"
#include "ng_philox.h"

namespace nestgpu
{

namespace
{

constexpr int GEN_BS = 256;                 // threads per workgroup
constexpr int SPT = 8;                      // sources per thread
constexpr uint32_t SEG = GEN_BS * SPT;      // sources per segment
constexpr int QPT = 4;                      // fixed_indegree draws per thread

// the words of one stream with the last Philox block kept
struct WordCache
{
  uint32_t k0, k1;
  uint64_t blk = ~uint64_t( 0 );
  uint32_t v[ 4 ];
  __device__ uint32_t
  word( uint64_t n )
  {
    const uint64_t b = n >> 2;
    if ( b != blk )
    {
      philox_block( k0, k1, b, v );
      blk = b;
    }
    const uint32_t i = 3u - static_cast< uint32_t >( n & 3u );
    return i == 0 ? v[ 0 ] : i == 1 ? v[ 1 ] : i == 2 ? v[ 2 ] : v[ 3 ];
  }
};

// exclusive prefix of x over the workgroup; returns the total
__device__ uint32_t
wg_scan( uint32_t x, uint32_t* ex )
{
  __shared__ uint32_t wsum[ GEN_BS / 64 ];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t inc = x;
  for ( int o = 1; o < 64; o <<= 1 )
  {
    const uint32_t y = __shfl_up( inc, o, 64 );
    if ( lane >= static_cast< uint32_t >( o ) )
      inc += y;
  }
  if ( lane == 63 )
    wsum[ w ] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for ( uint32_t i = 0; i < GEN_BS / 64; ++i )
  {
    before += i < w ? wsum[ i ] : 0u;
    total += wsum[ i ];
  }
  __syncthreads();
  *ex = before + inc - x;
  return total;
}

template < bool EMIT >
__global__ void __launch_bounds__( GEN_BS )
  k_bernoulli( const uint32_t* pre, uint32_t n_src, const GenRow* rows, const GenStream* streams, uint32_t n_chunks,
    double p, uint32_t* counts, const uint64_t* offsets, uint32_t* out_src, uint32_t* out_tgt )
{
  const uint64_t seg = blockIdx.x;
  const GenRow R = rows[ seg / n_chunks ];
  const GenStream S = streams[ R.vp ];
  const uint32_t j0 = static_cast< uint32_t >( seg % n_chunks ) * SEG + threadIdx.x * SPT;
  WordCache wc;
  wc.k0 = S.k0;
  wc.k1 = S.k1;
  uint32_t hits = 0;
#pragma unroll
  for ( int q = 0; q < SPT; ++q )
  {
    const uint32_t j = j0 + q;
    if ( j >= n_src or static_cast< int32_t >( j ) == R.aut )
      continue;
    // BernoulliBuilder: the autapse is skipped before the draw, so later sources move up one
    const uint64_t d = R.draw0 + j - ( R.aut >= 0 and j > static_cast< uint32_t >( R.aut ) ? 1u : 0u );
    const uint64_t n = S.pos + 2 * d;
    const uint32_t lo = wc.word( n );
    const uint32_t hi = wc.word( n + 1 );
    if ( philox_canonical( lo, hi ) < p )  // rng->drand() >= p_: no connection
      hits |= 1u << q;
  }
  uint32_t ex;
  const uint32_t total = wg_scan( static_cast< uint32_t >( __popc( hits ) ), &ex );
  if ( not EMIT )
  {
    if ( threadIdx.x == 0 )
      counts[ seg ] = total;
    return;
  }
  uint64_t o = offsets[ seg ] + ex;
#pragma unroll
  for ( int q = 0; q < SPT; ++q )
    if ( hits & ( 1u << q ) )
    {
      out_src[ o ] = pre[ j0 + q ];
      out_tgt[ o ] = R.tgt;
      ++o;
    }
}

// rejected words among n0 .. n0 + n_words - 1: one aligned block of four words per thread
__global__ void __launch_bounds__( GEN_BS )
  k_rejected( GenStream S, uint64_t n0, uint64_t n_words, uint32_t range, uint64_t* rej, uint32_t* n_rej, uint32_t cap )
{
  const uint64_t b = ( n0 >> 2 ) + static_cast< uint64_t >( blockIdx.x ) * GEN_BS + threadIdx.x;
  const uint64_t end = n0 + n_words;
  if ( b * 4 >= end )
    return;
  uint32_t v[ 4 ];
  philox_block( S.k0, S.k1, b, v );
#pragma unroll
  for ( int i = 0; i < 4; ++i )
  {
    const uint64_t n = b * 4 + static_cast< uint64_t >( i );
    if ( n < n0 or n >= end )
      continue;
    if ( not lemire_accepts( v[ 3 - i ], range ) )
    {
      const uint32_t k = atomicAdd( n_rej, 1u );
      if ( k < cap )
        rej[ k ] = n;
    }
  }
}

__global__ void __launch_bounds__( GEN_BS )
  k_fixed_indegree( GenStream S, uint64_t n_draws, uint32_t K, const uint32_t* pre, uint32_t range,
    const uint32_t* row_tgt, const uint64_t* rk, uint32_t n_rk, uint32_t* out_src, uint32_t* out_tgt )
{
  const uint64_t q0 = ( static_cast< uint64_t >( blockIdx.x ) * GEN_BS + threadIdx.x ) * QPT;
  if ( q0 >= n_draws )
    return;
  WordCache wc;
  wc.k0 = S.k0;
  wc.k1 = S.k1;
  // m: the rejected words before draw q0's word (rk[i] = r_i - pos - i is non-decreasing)
  uint32_t lo = 0, hi = n_rk;
  while ( lo < hi )
  {
    const uint32_t mid = ( lo + hi ) >> 1;
    if ( rk[ mid ] <= q0 )
      lo = mid + 1;
    else
      hi = mid;
  }
  uint32_t m = lo;
#pragma unroll
  for ( int i = 0; i < QPT; ++i )
  {
    const uint64_t q = q0 + static_cast< uint64_t >( i );
    if ( q >= n_draws )
      break;
    while ( m < n_rk and rk[ m ] <= q )
      ++m;
    const uint32_t w = wc.word( S.pos + q + m );
    out_src[ q ] = pre[ lemire_value( w, range ) ];
    out_tgt[ q ] = row_tgt[ q / K ];
  }
}

// Device memory of one call: stream-ordered allocations on a non-blocking stream of the
// calling thread, so that a call neither waits for nor holds up simulations that other
// threads run on the device (hipMalloc / hipFree synchronize the whole device).
struct CallMem
{
  hipStream_t st = nullptr;
  std::vector< void* > ptrs;
  explicit CallMem( int device )
  {
    thread_local hipStream_t streams[ 64 ] = {};
    if ( device < 0 or device >= 64 )
      throw std::runtime_error( "conngen: device index out of range" );
    if ( not streams[ device ] )
      hip_check( hipStreamCreateWithFlags( &streams[ device ], hipStreamNonBlocking ), "conngen: stream" );
    st = streams[ device ];
  }
  ~CallMem()
  {
    for ( void* p : ptrs )
      hipFreeAsync( p, st );
    hipStreamSynchronize( st );
  }
  template < typename T >
  T*
  alloc( size_t n )
  {
    void* p = nullptr;
    hip_check( hipMallocAsync( &p, std::max< size_t >( n, 1 ) * sizeof( T ), st ), "conngen: hipMallocAsync" );
    ptrs.push_back( p );
    return static_cast< T* >( p );
  }
  template < typename T >
  T*
  upload( const std::vector< T >& v )
  {
    T* p = alloc< T >( v.size() );
    if ( not v.empty() )
      hip_check( hipMemcpyAsync( p, v.data(), v.size() * sizeof( T ), hipMemcpyHostToDevice, st ), "conngen: upload" );
    return p;
  }
  // queued copy to host; valid after sync()
  template < typename T >
  void
  download( std::vector< T >& v, const T* p, size_t n )
  {
    v.resize( n );
    if ( n )
      hip_check( hipMemcpyAsync( v.data(), p, n * sizeof( T ), hipMemcpyDeviceToHost, st ), "conngen: download" );
  }
  void
  sync()
  {
    hip_check( hipStreamSynchronize( st ), "conngen: synchronize" );
  }
};

}  // namespace

void
gen_pairwise_bernoulli( int device, const std::vector< uint32_t >& pre, const std::vector< GenRow >& rows,
  const std::vector< GenStream >& streams, double p, std::vector< std::vector< uint32_t > >& src,
  std::vector< std::vector< uint32_t > >& tgt )
{
  hip_check( hipSetDevice( device ), "hipSetDevice" );
  const uint32_t n_src = static_cast< uint32_t >( pre.size() );
  const uint32_t n_chunks = ( n_src + SEG - 1 ) / SEG;
  const uint64_t n_seg = static_cast< uint64_t >( rows.size() ) * n_chunks;
  for ( auto& v : src )
    v.clear();
  for ( auto& v : tgt )
    v.clear();
  if ( n_seg == 0 )
    return;
  if ( n_seg > 0x7fffffffull )
    throw std::runtime_error( "NotImplemented: pairwise_bernoulli on the device: too many (target, source) segments" );
  CallMem M( device );
  const uint32_t* d_pre = M.upload( pre );
  const GenRow* d_rows = M.upload( rows );
  const GenStream* d_st = M.upload( streams );
  uint32_t* d_cnt = M.alloc< uint32_t >( n_seg );
  const uint32_t grid = static_cast< uint32_t >( n_seg );
  k_bernoulli< false ><<< grid, GEN_BS, 0, M.st >>>( d_pre, n_src, d_rows, d_st, n_chunks, p, d_cnt, nullptr, nullptr,
    nullptr );
  hip_check( hipGetLastError(), "conngen: count launch" );
  std::vector< uint32_t > cnt;
  M.download( cnt, d_cnt, n_seg );
  M.sync();
  std::vector< uint64_t > off( n_seg );
  uint64_t total = 0;
  for ( uint64_t s = 0; s < n_seg; ++s )
  {
    off[ s ] = total;
    total += cnt[ s ];
  }
  const uint64_t* d_off = M.upload( off );
  uint32_t* d_src = M.alloc< uint32_t >( total );
  uint32_t* d_tgt = M.alloc< uint32_t >( total );
  k_bernoulli< true ><<< grid, GEN_BS, 0, M.st >>>( d_pre, n_src, d_rows, d_st, n_chunks, p, nullptr, d_off, d_src, d_tgt );
  hip_check( hipGetLastError(), "conngen: emit launch" );
  std::vector< uint32_t > all_src, all_tgt;
  M.download( all_src, d_src, total );
  M.download( all_tgt, d_tgt, total );
  M.sync();
  // rows are grouped by VP in builder order: each VP's connections are contiguous
  uint64_t at = 0;
  for ( size_t r = 0; r < rows.size(); )
  {
    const uint32_t vp = rows[ r ].vp;
    uint64_t n = 0;
    size_t r1 = r;
    for ( ; r1 < rows.size() and rows[ r1 ].vp == vp; ++r1 )
      for ( uint32_t c = 0; c < n_chunks; ++c )
        n += cnt[ r1 * n_chunks + c ];
    src[ vp ].insert( src[ vp ].end(), all_src.begin() + at, all_src.begin() + at + n );
    tgt[ vp ].insert( tgt[ vp ].end(), all_tgt.begin() + at, all_tgt.begin() + at + n );
    at += n;
    r = r1;
  }
}

void
gen_fixed_indegree( int device, const std::vector< uint32_t >& pre,
  const std::vector< std::vector< uint32_t > >& row_tgt, uint32_t K, std::vector< GenStream >& streams,
  std::vector< std::vector< uint32_t > >& src, std::vector< std::vector< uint32_t > >& tgt )
{
  hip_check( hipSetDevice( device ), "hipSetDevice" );
  const uint32_t range = static_cast< uint32_t >( pre.size() );
  if ( pre.size() > 0xffffffffull or range == 0 )
    throw std::runtime_error( "NotImplemented: fixed_indegree on the device needs 1 to 2^32 - 1 sources" );
  // the share of words Lemire's method rejects for this range: threshold / 2^32
  const double reject_rate = static_cast< double >( ( 0u - range ) % range ) / 4294967296.0;
  const size_t n_vp = streams.size();
  std::vector< uint64_t > Q( n_vp ), slack( n_vp );
  for ( size_t vp = 0; vp < n_vp; ++vp )
  {
    src[ vp ].clear();
    tgt[ vp ].clear();
    Q[ vp ] = static_cast< uint64_t >( row_tgt[ vp ].size() ) * K;
    slack[ vp ] = 1024 + static_cast< uint64_t >( 2.0 * reject_rate * static_cast< double >( Q[ vp ] ) ) + Q[ vp ] / 4096;
  }
  CallMem M( device );
  const uint32_t* d_pre = M.upload( pre );
  // 1. the rejected words over a window each VP's draws reach: Q words plus room for the
  //    rejections; a window holding more rejections than its room is searched again, larger
  std::vector< std::vector< uint64_t > > rej( n_vp );
  std::vector< bool > done( n_vp, false );
  for ( size_t vp = 0; vp < n_vp; ++vp )
    done[ vp ] = Q[ vp ] == 0;
  while ( true )
  {
    std::vector< size_t > todo;
    for ( size_t vp = 0; vp < n_vp; ++vp )
      if ( not done[ vp ] )
        todo.push_back( vp );
    if ( todo.empty() )
      break;
    uint32_t* d_n = M.alloc< uint32_t >( n_vp );
    hip_check( hipMemsetAsync( d_n, 0, n_vp * sizeof( uint32_t ), M.st ), "conngen: clear" );
    std::vector< uint64_t* > d_rej( n_vp, nullptr );
    std::vector< uint32_t > cap( n_vp, 0 );
    for ( size_t vp : todo )
    {
      const GenStream& S = streams[ vp ];
      const uint64_t n_words = Q[ vp ] + slack[ vp ];
      cap[ vp ] = static_cast< uint32_t >( std::min< uint64_t >( slack[ vp ] + 1, 0x7fffffffull ) );
      d_rej[ vp ] = M.alloc< uint64_t >( cap[ vp ] );
      const uint64_t blocks = ( ( S.pos + n_words + 3 ) >> 2 ) - ( S.pos >> 2 );
      const uint64_t grid = ( blocks + GEN_BS - 1 ) / GEN_BS;
      if ( grid > 0x7fffffffull )
        throw std::runtime_error( "NotImplemented: fixed_indegree on the device: too many draws in one call" );
      k_rejected<<< static_cast< uint32_t >( grid ), GEN_BS, 0, M.st >>>( S, S.pos, n_words, range, d_rej[ vp ],
        d_n + vp, cap[ vp ] );
      hip_check( hipGetLastError(), "conngen: rejection launch" );
    }
    std::vector< uint32_t > n_rej;
    M.download( n_rej, d_n, n_vp );
    M.sync();
    for ( size_t vp : todo )
    {
      // the window holds every word the draws take when it holds at most `slack` rejections
      if ( n_rej[ vp ] <= slack[ vp ] and n_rej[ vp ] <= cap[ vp ] )
      {
        M.download( rej[ vp ], d_rej[ vp ], n_rej[ vp ] );
        done[ vp ] = true;
      }
      else
        slack[ vp ] *= 4;
    }
    M.sync();
  }
  // 2. every draw's word, all VPs queued together
  std::vector< std::vector< uint64_t > > rk( n_vp );
  for ( size_t vp = 0; vp < n_vp; ++vp )
  {
    if ( Q[ vp ] == 0 )
      continue;
    const GenStream& S = streams[ vp ];
    std::sort( rej[ vp ].begin(), rej[ vp ].end() );
    rk[ vp ].resize( rej[ vp ].size() );
    for ( size_t i = 0; i < rej[ vp ].size(); ++i )
      rk[ vp ][ i ] = rej[ vp ][ i ] - S.pos - i;
    const uint64_t* d_rk = M.upload( rk[ vp ] );
    const uint32_t* d_rows = M.upload( row_tgt[ vp ] );
    uint32_t* d_src = M.alloc< uint32_t >( Q[ vp ] );
    uint32_t* d_tgt = M.alloc< uint32_t >( Q[ vp ] );
    const uint64_t threads = ( Q[ vp ] + QPT - 1 ) / QPT;
    const uint64_t grid = ( threads + GEN_BS - 1 ) / GEN_BS;
    if ( grid > 0x7fffffffull )
      throw std::runtime_error( "NotImplemented: fixed_indegree on the device: too many draws in one call" );
    k_fixed_indegree<<< static_cast< uint32_t >( grid ), GEN_BS, 0, M.st >>>(
      S, Q[ vp ], K, d_pre, range, d_rows, d_rk, static_cast< uint32_t >( rk[ vp ].size() ), d_src, d_tgt );
    hip_check( hipGetLastError(), "conngen: draw launch" );
    M.download( src[ vp ], d_src, Q[ vp ] );
    M.download( tgt[ vp ], d_tgt, Q[ vp ] );
  }
  M.sync();
  for ( size_t vp = 0; vp < n_vp; ++vp )
    if ( Q[ vp ] )
    {
      // the stream continues after the last draw's word
      const uint64_t m_last =
        static_cast< uint64_t >( std::upper_bound( rk[ vp ].begin(), rk[ vp ].end(), Q[ vp ] - 1 ) - rk[ vp ].begin() );
      streams[ vp ].pos += Q[ vp ] + m_last;
    }
}

}  // namespace nestgpu
