// ng_kernels.hip -- gfx950 kernels for one min_delay slice of NEST's simulation step.
//
// Per slice s of a chunk, for every network instance in the launch:
//   k_expand   deliver the previous slice's spikes: walk each spike's out-row (CSR by
//              source, delay-sorted) and stage one record per synaptic event in the
//              target's slot range (nestkernel/event_delivery_manager.cpp:599-776,
//              connector_base.h:399-428)
//   k_big      targets with many events: sort their records in LDS and add them to the
//              ring buffers in receive order (ring_buffer.h:157-161,405-408)
//   k_update_* neuron update over [from, to) (models/*.cpp update()), fused with the
//              per-target reduction of small event lists and the spike_generator drive
//   k_compact  ordered compaction of spike flags into the receive-order spike list
//              (event_delivery_manager_impl.h:103-118, event_delivery_manager.cpp:470-497)
//
// Determinism: FP64 additions into one ring-buffer cell happen in NEST's order -- by
// delivery batch, drive after the batch of the same slice, receive order within a batch
// and connector-run order within a spike. The device sorts each target's records by
// (spike index, row position) before adding, so no result depends on atomic order.
// Built with -ffp-contract=off: every multiply and add is rounded as in the reference.

#include <hip/hip_runtime.h>

#include "ng_aeif.h"
#include "ng_device.h"
#include "ng_engine.h"

namespace nestgpu
{

constexpr int BS = 256;
constexpr int BIG_LDS = 2048;  // records sorted in LDS by one workgroup

__device__ inline uint64_t
rec_key( const Rec& r )
{
  return ( static_cast< uint64_t >( r.k ) << 32 ) | r.pr;
}

__device__ inline int
model_of( const NetDev& d, uint32_t dev )
{
  int m = 0;
  while ( m + 1 < N_MODELS and dev >= d.model_begin[ m + 1 ] )
    ++m;
  return m;
}

// handle(SpikeEvent&) of the three models: channel choice and value added.
__device__ inline void
rb_add( const NetDev& d, int m, uint32_t j, int64_t abs_step, double w, double mult )
{
  const uint32_t nm = d.model_begin[ m + 1 ] - d.model_begin[ m ];
  const int C = model_channels( m );
  const int64_t slot = abs_step % d.L;
  double* base = d.rb[ m ] + static_cast< size_t >( slot ) * C * nm;
  if ( m == M_IAF_PSC_DELTA )
    base[ j ] += w * mult;  // iaf_psc_delta::handle
  else if ( m == M_IAF_PSC_EXP )
  {
    const double s = w * mult;  // iaf_psc_exp::handle: SYN_IN = 0, SYN_EX = 1
    base[ ( s > 0 ? 1 : 0 ) * static_cast< size_t >( nm ) + j ] += s;
  }
  else
  {
    if ( w > 0.0 )  // aeif_cond_exp::handle
      base[ j ] += w * mult;
    else
      base[ static_cast< size_t >( nm ) + j ] += -w * mult;
  }
}

// ------------------------------------------------------------------ expand
__global__ void __launch_bounds__( BS ) k_expand( const NetDev* __restrict__ nets, int n_nets, int s )
{
  for ( int i = 0; i < n_nets; ++i )
  {
    const NetDev& d = nets[ i ];
    const SlicePhase ph = slice_phase( d, s );
    if ( not ph.deliver )
      continue;
    const uint32_t K = d.ctr[ 0 ];
    const uint64_t items = static_cast< uint64_t >( K ) * d.cmax;
    const uint32_t tmask = ( d.tgt_bits >= 32 ) ? 0xffffffffu : ( ( 1u << d.tgt_bits ) - 1u );
    for ( uint64_t it = blockIdx.x; it < items; it += gridDim.x )
    {
      const uint32_t k = static_cast< uint32_t >( it / d.cmax );
      const uint32_t c = static_cast< uint32_t >( it - static_cast< uint64_t >( k ) * d.cmax );
      const uint32_t src = d.spk[ k ];
      const uint32_t lag = d.spk_lag[ k ];
      const uint64_t r0 = d.row_ptr[ src ];
      const uint64_t len = d.row_ptr[ src + 1 ] - r0;
      const uint64_t pos = static_cast< uint64_t >( c ) * BS + threadIdx.x;
      if ( pos >= len )
        continue;
      const uint64_t g = r0 + pos;
      const uint32_t word = d.tgt_d[ g ];
      const double w = d.wgt[ g ];
      const uint32_t t = word & tmask;
      const uint32_t rel = lag + ( d.tgt_bits >= 32 ? 0u : ( word >> d.tgt_bits ) );
      const uint32_t p = atomicAdd( &d.cnt[ t ], 1u );
      if ( p == static_cast< uint32_t >( d.small_max ) )
      {
        const uint32_t b = atomicAdd( &d.ctr[ 1 ], 1u );
        d.big[ b ] = t;
      }
      const uint64_t ib = d.in_ptr[ t ];
      const uint64_t cap = ( d.in_ptr[ t + 1 ] - ib ) * static_cast< uint64_t >( d.s_cap );
      if ( p < cap )
      {
        Rec r;
        r.k = k;
        r.pr = ( static_cast< uint32_t >( pos ) << d.slot_bits ) | rel;
        r.w = w;
        d.rec[ ib * d.s_cap + p ] = r;
      }
      else
        atomicOr( &d.ctr[ 2 ], 1u );
    }
  }
}

// ------------------------------------------------------------------ big targets
// Bitonic sort of up to BIG_LDS records in LDS; larger lists are merge-sorted in global
// memory through rec_tmp (one region per workgroup, sized by the largest capacity).
__device__ void
wg_sort_lds( uint64_t* key, uint32_t* idx, int n_pow2 )
{
  for ( int kk = 2; kk <= n_pow2; kk <<= 1 )
    for ( int jj = kk >> 1; jj > 0; jj >>= 1 )
    {
      for ( int i = threadIdx.x; i < n_pow2; i += blockDim.x )
      {
        const int ixj = i ^ jj;
        if ( ixj > i )
        {
          const bool up = ( i & kk ) == 0;
          const uint64_t a = key[ i ], b = key[ ixj ];
          if ( ( a > b ) == up )
          {
            key[ i ] = b;
            key[ ixj ] = a;
            const uint32_t t = idx[ i ];
            idx[ i ] = idx[ ixj ];
            idx[ ixj ] = t;
          }
        }
      }
      __syncthreads();
    }
}

__global__ void __launch_bounds__( BS ) k_big( const NetDev* __restrict__ nets, int n_nets, int s )
{
  __shared__ uint64_t key[ BIG_LDS ];
  __shared__ uint32_t idx[ BIG_LDS ];
  for ( int i = 0; i < n_nets; ++i )
  {
    const NetDev& d = nets[ i ];
    const SlicePhase ph = slice_phase( d, s );
    if ( not ph.deliver )
      continue;
    const uint32_t nb = d.ctr[ 1 ];
    const uint32_t smask = ( 1u << d.slot_bits ) - 1u;
    for ( uint32_t b = blockIdx.x; b < nb; b += gridDim.x )
    {
      const uint32_t t = d.big[ b ];
      const uint64_t ib = d.in_ptr[ t ];
      const uint64_t cap = ( d.in_ptr[ t + 1 ] - ib ) * static_cast< uint64_t >( d.s_cap );
      uint64_t n = d.cnt[ t ];
      if ( n > cap )
        n = cap;  // overflow is flagged by k_expand; the run reports it
      Rec* R = d.rec + ib * d.s_cap;
      const int m = model_of( d, t );
      const uint32_t j = t - d.model_begin[ m ];
      const int C = model_channels( m );
      const int n_cells = d.max_delay * C;
      if ( n <= BIG_LDS )
      {
        int p2 = 1;
        while ( p2 < static_cast< int >( n ) )
          p2 <<= 1;
        for ( int q = threadIdx.x; q < p2; q += blockDim.x )
        {
          key[ q ] = q < static_cast< int >( n ) ? rec_key( R[ q ] ) : ~0ull;
          idx[ q ] = q;
        }
        __syncthreads();
        wg_sort_lds( key, idx, p2 );
        // each cell's adds are done by one thread, in key order
        for ( int cell = threadIdx.x; cell < n_cells; cell += blockDim.x )
        {
          for ( uint64_t q = 0; q < n; ++q )
          {
            const Rec r = R[ idx[ q ] ];
            const int rel = static_cast< int >( r.pr & smask );
            int ch = 0;
            if ( m == M_IAF_PSC_EXP )
              ch = r.w > 0 ? 1 : 0;
            else if ( m == M_AEIF_COND_EXP )
              ch = r.w > 0.0 ? 0 : 1;
            if ( rel * C + ch != cell )
              continue;
            rb_add( d, m, j, ph.origin + rel, r.w, 1.0 );
          }
        }
        __syncthreads();
      }
      else
      {
        // merge sort in global memory: LDS-sorted runs of BIG_LDS, then pairwise merges
        Rec* A = R;
        Rec* B = d.rec_tmp + static_cast< uint64_t >( blockIdx.x ) * d.max_cap;
        for ( uint64_t r0 = 0; r0 < n; r0 += BIG_LDS )
        {
          const int len = static_cast< int >( n - r0 < BIG_LDS ? n - r0 : BIG_LDS );
          int p2 = 1;
          while ( p2 < len )
            p2 <<= 1;
          for ( int q = threadIdx.x; q < p2; q += blockDim.x )
          {
            key[ q ] = q < len ? rec_key( A[ r0 + q ] ) : ~0ull;
            idx[ q ] = q;
          }
          __syncthreads();
          wg_sort_lds( key, idx, p2 );
          for ( int q = threadIdx.x; q < len; q += blockDim.x )
            B[ r0 + q ] = A[ r0 + idx[ q ] ];
          __syncthreads();
        }
        // B holds sorted runs; merge runs of width w into the other buffer
        Rec* src = B;
        Rec* dst = A;
        for ( uint64_t w = BIG_LDS; w < n; w <<= 1 )
        {
          for ( uint64_t o = threadIdx.x; o < n; o += blockDim.x )
          {
            // output position o belongs to the pair starting at lo
            const uint64_t lo = ( o / ( 2 * w ) ) * ( 2 * w );
            const uint64_t mid = lo + w < n ? lo + w : n;
            const uint64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            const uint64_t rank = o - lo;
            // co-rank: number a taken from [lo, mid) among the first rank outputs
            uint64_t a_lo = rank > ( hi - mid ) ? rank - ( hi - mid ) : 0;
            uint64_t a_hi = rank < ( mid - lo ) ? rank : ( mid - lo );
            while ( a_lo < a_hi )
            {
              const uint64_t a = ( a_lo + a_hi ) / 2;
              const uint64_t bb = rank - a;
              // take more from A-run if A[a] < B[bb-1]
              if ( bb > 0 and a < mid - lo and rec_key( src[ lo + a ] ) < rec_key( src[ mid + bb - 1 ] ) )
                a_lo = a + 1;
              else
                a_hi = a;
            }
            const uint64_t a = a_lo, bb = rank - a;
            bool take_a;
            if ( a >= mid - lo )
              take_a = false;
            else if ( bb >= hi - mid )
              take_a = true;
            else
              take_a = rec_key( src[ lo + a ] ) < rec_key( src[ mid + bb ] );
            dst[ o ] = take_a ? src[ lo + a ] : src[ mid + bb ];
          }
          __syncthreads();
          Rec* tmp = src;
          src = dst;
          dst = tmp;
        }
        for ( int cell = threadIdx.x; cell < n_cells; cell += blockDim.x )
        {
          for ( uint64_t q = 0; q < n; ++q )
          {
            const Rec r = src[ q ];
            const int rel = static_cast< int >( r.pr & smask );
            int ch = 0;
            if ( m == M_IAF_PSC_EXP )
              ch = r.w > 0 ? 1 : 0;
            else if ( m == M_AEIF_COND_EXP )
              ch = r.w > 0.0 ? 0 : 1;
            if ( rel * C + ch != cell )
              continue;
            rb_add( d, m, j, ph.origin + rel, r.w, 1.0 );
          }
        }
        __syncthreads();
      }
      if ( threadIdx.x == 0 )
        d.cnt[ t ] = 0;
      __syncthreads();
    }
  }
}

// Small per-target lists: the update thread adds its records in key order by selection.
__device__ inline void
reduce_small( const NetDev& d, int m, uint32_t dev, uint32_t j, int64_t origin )
{
  const uint32_t c = d.cnt[ dev ];
  if ( c == 0 )
    return;
  d.cnt[ dev ] = 0;
  if ( c > static_cast< uint32_t >( d.small_max ) )
    return;  // handled by k_big (cannot happen: k_big resets the count)
  const uint64_t ib = d.in_ptr[ dev ];
  const uint64_t cap = ( d.in_ptr[ dev + 1 ] - ib ) * static_cast< uint64_t >( d.s_cap );
  const uint32_t n = c < cap ? c : static_cast< uint32_t >( cap );
  const Rec* R = d.rec + ib * d.s_cap;
  const uint32_t smask = ( 1u << d.slot_bits ) - 1u;
  uint64_t last = 0;
  for ( uint32_t it = 0; it < n; ++it )
  {
    uint64_t best = ~0ull;
    uint32_t bi = 0;
    for ( uint32_t q = 0; q < n; ++q )
    {
      const uint64_t kq = rec_key( R[ q ] );
      if ( ( it == 0 or kq > last ) and kq < best )
      {
        best = kq;
        bi = q;
      }
    }
    last = best;
    const Rec r = R[ bi ];
    rb_add( d, m, j, origin + static_cast< int64_t >( r.pr & smask ), r.w, 1.0 );
  }
}

// Poisson count for seeded drive: splitmix64 of (seed, stream, step), inverse CDF
__host__ __device__ inline uint32_t
poisson_count( const PoissonTable& T, uint64_t stream, int64_t step )
{
  uint64_t z = T.seed + stream * 0x9E3779B97F4A7C15ull + static_cast< uint64_t >( step ) * 0xD1B54A32D192ED03ull;
  z = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
  z = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
  z = z ^ ( z >> 31 );
  const double u = static_cast< double >( z >> 11 ) * ( 1.0 / 9007199254740992.0 );
  uint32_t k = 0;
  while ( static_cast< int32_t >( k ) < T.kmax and u >= T.cdf[ k ] )
    ++k;
  return k;
}

__device__ inline void
drive_adds( const NetDev& d, const ParamSet& ps, int m, uint32_t dev, uint32_t j, int64_t origin, int from, int to )
{
  if ( ps.drive_mode == 0 )
    return;
  const uint32_t col = d.drive_idx[ dev ];
  if ( col == 0xffffffffu )
    return;
  for ( int lag = from; lag < to; ++lag )
  {
    const int64_t e = origin + lag;  // emission step: spike stamp e+1
    if ( e < ps.drive_t0 or e >= ps.drive_t0 + ps.drive_steps )
      continue;
    uint32_t c;
    if ( ps.drive_mode == 1 )
      c = d.drive_counts[ static_cast< uint64_t >( e - ps.drive_t0 ) * d.drive_cols + col ];
    else
      c = poisson_count( d.ptables[ ps.drive_table ], col, e );
    if ( c )
      rb_add( d, m, j, e + ps.drive_d, ps.drive_w, static_cast< double >( c ) );
  }
}

__device__ inline void
write_mask( const NetDev& d, uint32_t dev, const uint64_t* bits )
{
  const uint32_t o = d.dev2ord ? d.dev2ord[ dev ] : dev;
  for ( int w = 0; w < d.nw; ++w )
    d.mask[ static_cast< uint64_t >( o ) * d.nw + w ] = bits[ w ];
}

// ------------------------------------------------------------------ iaf_psc_delta
// models/iaf_psc_delta.cpp:272-331 (refractory_input = false)
__global__ void __launch_bounds__( BS ) k_update_iaf_psc_delta( const NetDev* __restrict__ nets, int n_nets, int s )
{
  constexpr int m = M_IAF_PSC_DELTA;
  for ( int i = 0; i < n_nets; ++i )
  {
    const NetDev& d = nets[ i ];
    const SlicePhase ph = slice_phase( d, s );
    if ( not ph.active )
      continue;
    const uint32_t b0 = d.model_begin[ m ];
    const uint32_t nm = d.model_begin[ m + 1 ] - b0;
    for ( uint32_t j = blockIdx.x * BS + threadIdx.x; j < nm; j += gridDim.x * BS )
    {
      const uint32_t dev = b0 + j;
      if ( ph.deliver )
        reduce_small( d, m, dev, j, ph.origin );
      const ParamSet& ps = d.psets[ d.pidx[ m ][ j ] ];
      const double P30 = ps.p[ 0 ], P33 = ps.p[ 1 ], y0I = ps.p[ 2 ], V_th = ps.p[ 3 ], V_reset = ps.p[ 4 ],
                   V_min = ps.p[ 5 ];
      double y3 = d.st[ m ][ 0 ][ j ];
      int32_t r = d.r[ m ][ j ];
      uint64_t bits[ 4 ] = { 0, 0, 0, 0 };
      for ( int lag = ph.from; lag < ph.to; ++lag )
      {
        const int64_t slot = ( ph.origin + lag ) % d.L;
        double* cell = d.rb[ m ] + static_cast< size_t >( slot ) * nm + j;
        const double in = *cell;
        *cell = 0.0;
        if ( r == 0 )
        {
          y3 = P30 * y0I + P33 * y3 + in;
          y3 = ( y3 < V_min ? V_min : y3 );
        }
        else
          --r;
        if ( y3 >= V_th )
        {
          r = ps.refractory_counts;
          y3 = V_reset;
          bits[ lag >> 6 ] |= 1ull << ( lag & 63 );
        }
      }
      drive_adds( d, ps, m, dev, j, ph.origin, ph.from, ph.to );
      d.st[ m ][ 0 ][ j ] = y3;
      d.r[ m ][ j ] = r;
      write_mask( d, dev, bits );
    }
  }
}

// ------------------------------------------------------------------ iaf_psc_exp
// models/iaf_psc_exp.cpp:288-349 (delta = 0: deterministic threshold)
__global__ void __launch_bounds__( BS ) k_update_iaf_psc_exp( const NetDev* __restrict__ nets, int n_nets, int s )
{
  constexpr int m = M_IAF_PSC_EXP;
  for ( int i = 0; i < n_nets; ++i )
  {
    const NetDev& d = nets[ i ];
    const SlicePhase ph = slice_phase( d, s );
    if ( not ph.active )
      continue;
    const uint32_t b0 = d.model_begin[ m ];
    const uint32_t nm = d.model_begin[ m + 1 ] - b0;
    for ( uint32_t j = blockIdx.x * BS + threadIdx.x; j < nm; j += gridDim.x * BS )
    {
      const uint32_t dev = b0 + j;
      if ( ph.deliver )
        reduce_small( d, m, dev, j, ph.origin );
      const ParamSet& ps = d.psets[ d.pidx[ m ][ j ] ];
      const double P22 = ps.p[ 0 ], P21ex = ps.p[ 1 ], P21in = ps.p[ 2 ], P20 = ps.p[ 3 ], P11ex = ps.p[ 4 ],
                   P11in = ps.p[ 5 ], IeI0 = ps.p[ 6 ], Theta = ps.p[ 7 ], V_reset = ps.p[ 8 ], exI1 = ps.p[ 9 ];
      double V = d.st[ m ][ 0 ][ j ];
      double iex = d.st[ m ][ 1 ][ j ];
      double iin = d.st[ m ][ 2 ][ j ];
      int32_t r = d.r[ m ][ j ];
      uint64_t bits[ 4 ] = { 0, 0, 0, 0 };
      for ( int lag = ph.from; lag < ph.to; ++lag )
      {
        const int64_t slot = ( ph.origin + lag ) % d.L;
        double* cin = d.rb[ m ] + static_cast< size_t >( slot ) * 2 * nm + j;
        double* cex = cin + nm;
        if ( r == 0 )
          V = V * P22 + iex * P21ex + iin * P21in + IeI0 * P20;
        else
          --r;
        iex *= P11ex;
        iin *= P11in;
        iex += exI1;
        const double wex = *cex;
        const double win = *cin;
        iex += wex;
        iin += win;
        if ( V >= Theta )
        {
          r = ps.refractory_counts;
          V = V_reset;
          bits[ lag >> 6 ] |= 1ull << ( lag & 63 );
        }
        *cex = 0.0;
        *cin = 0.0;
      }
      drive_adds( d, ps, m, dev, j, ph.origin, ph.from, ph.to );
      d.st[ m ][ 0 ][ j ] = V;
      d.st[ m ][ 1 ][ j ] = iex;
      d.st[ m ][ 2 ][ j ] = iin;
      d.r[ m ][ j ] = r;
      write_mask( d, dev, bits );
    }
  }
}

// ------------------------------------------------------------------ aeif_cond_exp
// models/aeif_cond_exp.cpp:443-523
__global__ void __launch_bounds__( BS ) k_update_aeif_cond_exp( const NetDev* __restrict__ nets, int n_nets, int s )
{
  constexpr int m = M_AEIF_COND_EXP;
  for ( int i = 0; i < n_nets; ++i )
  {
    const NetDev& d = nets[ i ];
    const SlicePhase ph = slice_phase( d, s );
    if ( not ph.active )
      continue;
    const uint32_t b0 = d.model_begin[ m ];
    const uint32_t nm = d.model_begin[ m + 1 ] - b0;
    for ( uint32_t j = blockIdx.x * BS + threadIdx.x; j < nm; j += gridDim.x * BS )
    {
      const uint32_t dev = b0 + j;
      if ( ph.deliver )
        reduce_small( d, m, dev, j, ph.origin );
      const ParamSet& ps = d.psets[ d.pidx[ m ][ j ] ];
      AeifPar P;
      P.V_peak = ps.p[ 0 ];
      P.V_reset = ps.p[ 1 ];
      P.g_L = ps.p[ 2 ];
      P.C_m = ps.p[ 3 ];
      P.E_ex = ps.p[ 4 ];
      P.E_in = ps.p[ 5 ];
      P.E_L = ps.p[ 6 ];
      P.Delta_T = ps.p[ 7 ];
      P.tau_w = ps.p[ 8 ];
      P.a = ps.p[ 9 ];
      P.b = ps.p[ 10 ];
      P.V_th = ps.p[ 11 ];
      P.tau_ex = ps.p[ 12 ];
      P.tau_in = ps.p[ 13 ];
      P.I_e = ps.p[ 14 ];
      P.tol = ps.p[ 15 ];
      const double V_peak_eff = ps.p[ 16 ];
      const double hstep = ps.p[ 17 ];
      double y[ 4 ] = { d.st[ m ][ 0 ][ j ], d.st[ m ][ 1 ][ j ], d.st[ m ][ 2 ][ j ], d.st[ m ][ 3 ][ j ] };
      double hint = d.st[ m ][ 4 ][ j ];
      int32_t r = d.r[ m ][ j ];
      uint64_t bits[ 4 ] = { 0, 0, 0, 0 };
      for ( int lag = ph.from; lag < ph.to; ++lag )
      {
        double t = 0.0;
        int guard = 0;
        while ( t < hstep and guard < 1000000 )
        {
          ++guard;
          aeif_evolve( P, r > 0, &t, hstep, &hint, y );
          if ( y[ 0 ] < -1e3 or y[ 3 ] < -1e6 or y[ 3 ] > 1e6 )
          {
            atomicOr( &d.ctr[ 2 ], 2u );  // NumericalInstability
            t = hstep;
            break;
          }
          if ( r > 0 )
            y[ 0 ] = P.V_reset;
          else if ( y[ 0 ] >= V_peak_eff )
          {
            y[ 0 ] = P.V_reset;
            y[ 3 ] += P.b;
            r = ps.refractory_counts > 0 ? ps.refractory_counts + 1 : 0;
            if ( bits[ lag >> 6 ] & ( 1ull << ( lag & 63 ) ) )
              atomicOr( &d.ctr[ 2 ], 4u );  // second spike within one step
            bits[ lag >> 6 ] |= 1ull << ( lag & 63 );
          }
        }
        if ( r > 0 )
          --r;
        const int64_t slot = ( ph.origin + lag ) % d.L;
        double* cex = d.rb[ m ] + static_cast< size_t >( slot ) * 2 * nm + j;
        double* cinh = cex + nm;
        y[ 1 ] += *cex;
        y[ 2 ] += *cinh;
        *cex = 0.0;
        *cinh = 0.0;
      }
      drive_adds( d, ps, m, dev, j, ph.origin, ph.from, ph.to );
      for ( int q = 0; q < 4; ++q )
        d.st[ m ][ q ][ j ] = y[ q ];
      d.st[ m ][ 4 ][ j ] = hint;
      d.r[ m ][ j ] = r;
      write_mask( d, dev, bits );
    }
  }
}

// ------------------------------------------------------------------ compaction
// Decoupled look-back scan over order positions (one tile of BS per workgroup step).
// Tile state: (epoch << 32) | (flag << 30) | value, flag 1 = aggregate, 2 = inclusive.
__global__ void __launch_bounds__( BS ) k_compact( const NetDev* __restrict__ nets, int n_nets, int s )
{
  __shared__ uint32_t warp_sums[ BS / 64 ];
  __shared__ uint32_t s_excl;
  __shared__ uint32_t s_base;
  for ( int i = 0; i < n_nets; ++i )
  {
    const NetDev& d = nets[ i ];
    const SlicePhase ph = slice_phase( d, s );
    if ( not ph.active )
      continue;
    const uint32_t base = ph.deliver ? 0u : d.ctr[ 0 ];
    const uint32_t epoch = d.chunk_seq * 1024u + static_cast< uint32_t >( s );
    const uint32_t n_tiles = ( d.n + BS - 1 ) / BS;
    uint64_t* state = reinterpret_cast< uint64_t* >( d.scan_state );
    for ( uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x )
    {
      const uint32_t o = tile * BS + threadIdx.x;
      uint32_t cnt = 0;
      uint64_t words[ 4 ] = { 0, 0, 0, 0 };
      if ( o < d.n )
        for ( int w = 0; w < d.nw; ++w )
        {
          words[ w ] = d.mask[ static_cast< uint64_t >( o ) * d.nw + w ];
          cnt += __popcll( words[ w ] );
        }
      // block inclusive scan of cnt
      const int lane = threadIdx.x & 63;
      const int wid = threadIdx.x >> 6;
      uint32_t v = cnt;
      for ( int off = 1; off < 64; off <<= 1 )
      {
        const uint32_t u = __shfl_up( v, off, 64 );
        if ( lane >= off )
          v += u;
      }
      if ( lane == 63 )
        warp_sums[ wid ] = v;
      __syncthreads();
      uint32_t wpre = 0, total = 0;
      for ( int q = 0; q < BS / 64; ++q )
      {
        if ( q < wid )
          wpre += warp_sums[ q ];
        total += warp_sums[ q ];
      }
      const uint32_t local_excl = wpre + v - cnt;
      if ( threadIdx.x == 0 )
      {
        const uint64_t ep = static_cast< uint64_t >( epoch ) << 32;
        if ( tile == 0 )
        {
          __hip_atomic_store( &state[ tile ], ep | ( 2ull << 30 ) | total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT );
          s_excl = 0;
        }
        else
        {
          __hip_atomic_store( &state[ tile ], ep | ( 1ull << 30 ) | total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT );
          uint32_t excl = 0;
          int64_t jt = static_cast< int64_t >( tile ) - 1;
          long spins = 0;
          while ( jt >= 0 )
          {
            const uint64_t st = __hip_atomic_load( &state[ jt ], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT );
            const uint32_t f = static_cast< uint32_t >( ( st >> 30 ) & 3u );
            if ( ( st >> 32 ) != epoch or f == 0 )
            {
              if ( ++spins > ( 1l << 26 ) )
              {
                atomicOr( &d.ctr[ 2 ], 8u );  // look-back did not converge
                break;
              }
              continue;
            }
            excl += static_cast< uint32_t >( st & 0x3fffffffu );
            if ( f == 2 )
              break;
            --jt;
          }
          __hip_atomic_store(
            &state[ tile ], ep | ( 2ull << 30 ) | ( excl + total ), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT );
          s_excl = excl;
        }
        s_base = base;
      }
      __syncthreads();
      uint32_t off = s_base + s_excl + local_excl;
      if ( o < d.n and cnt )
      {
        const uint32_t dev = d.ord2dev ? d.ord2dev[ o ] : o;
        const unsigned long long lb = *d.log_n;
        for ( int w = 0; w < d.nw; ++w )
        {
          uint64_t x = words[ w ];
          while ( x )
          {
            const int bpos = __ffsll( static_cast< long long >( x ) ) - 1;
            x &= x - 1;
            const uint32_t lag = w * 64 + bpos;
            d.spk[ off ] = dev;
            d.spk_lag[ off ] = lag;
            const unsigned long long li = lb + ( off - s_base );
            if ( li < d.log_cap )
            {
              d.log_src[ li ] = dev;
              d.log_step[ li ] = ph.origin + lag + 1;
            }
            ++off;
          }
        }
      }
      if ( o < d.n and cnt )
        for ( int w = 0; w < d.nw; ++w )
          d.mask[ static_cast< uint64_t >( o ) * d.nw + w ] = 0;
      if ( tile == n_tiles - 1 and threadIdx.x == 0 )
      {
        d.ctr[ 0 ] = s_base + s_excl + total;
        d.ctr[ 4 ] = s_excl + total;  // spikes this part (for the log counter)
      }
      __syncthreads();
    }
  }
}

// Advances the spike-log counter and clears the big-target counter after compaction.
__global__ void
k_finish( const NetDev* __restrict__ nets, int n_nets, int s )
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if ( i >= n_nets )
    return;
  const NetDev& d = nets[ i ];
  const SlicePhase ph = slice_phase( d, s );
  if ( not ph.active )
    return;
  *d.log_n += d.ctr[ 4 ];
  d.ctr[ 4 ] = 0;
  d.ctr[ 1 ] = 0;
  if ( ph.deliver )
    d.ctr[ 3 ] = 0;
  d.ctr[ 3 ] += 1;  // parts in the current slice
  if ( ph.complete and d.n_vp > 1 and d.ctr[ 3 ] > 1 )
  {
    // A slice split across Run calls: NEST's registers hold each VP's spikes part by
    // part, so the receive order is (VP, part, node, lag). Restore it with a stable
    // counting pass by VP over the parts appended in order (rare; single thread).
    const uint32_t K = d.ctr[ 0 ];
    uint32_t* tsrc = d.spk_tmp;
    uint32_t* tlag = d.spk_tmp + K;
    for ( uint32_t q = 0; q < K; ++q )
    {
      tsrc[ q ] = d.spk[ q ];
      tlag[ q ] = d.spk_lag[ q ];
    }
    uint32_t o = 0;
    for ( int vp = 0; vp < d.n_vp; ++vp )
      for ( uint32_t q = 0; q < K; ++q )
      {
        const uint32_t dev = tsrc[ q ];
        const uint32_t ord = d.dev2ord ? d.dev2ord[ dev ] : dev;
        if ( ord >= d.vp_begin[ vp ] and ord < d.vp_begin[ vp + 1 ] )
        {
          d.spk[ o ] = dev;
          d.spk_lag[ o ] = tlag[ q ];
          ++o;
        }
      }
  }
}

// ------------------------------------------------------------------ launch wrappers
hipError_t
launch_slice( const NetDev* nets, int n_nets, int s, const bool* has_model, const LaunchCfg& cfg, hipStream_t st )
{
  k_expand<<< cfg.grid_expand, BS, 0, st >>>( nets, n_nets, s );
  k_big<<< cfg.grid_big, BS, 0, st >>>( nets, n_nets, s );
  if ( has_model[ M_IAF_PSC_DELTA ] )
    k_update_iaf_psc_delta<<< cfg.grid_update, BS, 0, st >>>( nets, n_nets, s );
  if ( has_model[ M_IAF_PSC_EXP ] )
    k_update_iaf_psc_exp<<< cfg.grid_update, BS, 0, st >>>( nets, n_nets, s );
  if ( has_model[ M_AEIF_COND_EXP ] )
    k_update_aeif_cond_exp<<< cfg.grid_update, BS, 0, st >>>( nets, n_nets, s );
  k_compact<<< cfg.grid_compact, BS, 0, st >>>( nets, n_nets, s );
  k_finish<<< ( n_nets + 63 ) / 64, 64, 0, st >>>( nets, n_nets, s );
  return hipGetLastError();
}

}  // namespace nestgpu
