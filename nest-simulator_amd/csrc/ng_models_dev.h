// ng_models_dev.h -- the device halves of the neuron models.
//
// For each model id registered with ModelManager (ng_model.h), DeviceNode< id > gives what
// nest::Node gives the reference's update loop (nestkernel/node.h:297,572):
//   - receive: handle( SpikeEvent& ) -- the ring-buffer channel an event of weight w (weight
//     times multiplicity) goes to, and the value added there;
//   - update: update( origin, from, to ) of one neuron over the lags [ph.from, ph.to) of the
//     slice at ph.origin: read its state (fields v0..v4 and r of its device index), consume and
//     clear its ring-buffer cells lag by lag, set the bit of every lag it spikes in, store the
//     multimeter samples due (store_sample at lag next_sample, then every mm_int lags) and
//     write the state back. The parameters are the ParamSet its host half's pre_run_hook made.
// The slice kernel dispatches on the model id (a compile-time id when a launch holds one
// model), so every model is inlined into the fused kernel. Included by ng_kernels.hip after
// the ring-buffer cell and sampling helpers it uses.
#pragma once

namespace nestgpu
{

template < int ID >
struct DeviceNode;

// The device model ids, in the order ModelManager registers them: a model with a DeviceNode
// specialization joins the slice kernel by its entry here. WIDE marks models whose launches
// run one workgroup per CU with up to 256 VGPRs; the others are held to 128 VGPRs, so that
// two workgroups share a CU, spilling what does not fit.
#define NG_DEVICE_MODELS( X ) X( M_IAF_PSC_DELTA ) X( M_IAF_PSC_EXP ) X( M_AEIF_COND_EXP ) X( M_IAF_PSC_ALPHA )
constexpr int N_DEVICE_MODELS = 4;
constexpr int ALL_MODELS = ( 1 << N_DEVICE_MODELS ) - 1;
static_assert( ALL_MODELS < MASK_STDP, "model bits must stay below the stdp bit" );

// the one model id of a launch mask with a single model bit, else -1
template < int MASK >
__device__ constexpr int
single_model()
{
  return ( MASK & ( MASK - 1 ) ) == 0 ? __builtin_ctz( MASK ) : -1;
}

template <>
struct DeviceNode< M_IAF_PSC_DELTA >
{
  static constexpr bool WIDE = false;
  // iaf_psc_delta::handle( SpikeEvent& ) (iaf_psc_delta.cpp:332-341): one channel, the weight
  __device__ __forceinline__ static void
  receive( double w, uint32_t& ch, double& v )
  {
    ch = 0;
    v = w;
  }
  template < bool WIN >
  __device__ __forceinline__ static void
  update( const NetDev& d, gref< const ParamSet > ps, uint32_t dev, uint32_t local, const Cells< WIN >& C,
    const SlicePhase& ph, uint64_t* bits, uint32_t mcol, int next_sample )
  {
    // models/iaf_psc_delta.cpp:272-331
    const double P30 = ps.p[ 0 ], P33 = ps.p[ 1 ], y0I = ps.p[ 2 ], V_th = ps.p[ 3 ], V_reset = ps.p[ 4 ],
                 V_min = ps.p[ 5 ];
    const bool refr_input = ps.p[ 6 ] != 0.0;
    const double h = ps.p[ 7 ], tau_m = ps.p[ 8 ];
    double y3 = glob( d.v[ 0 ] )[ dev ];
    double refr_buf = refr_input ? glob( d.v[ 1 ] )[ dev ] : 0.0;  // refr_spikes_buffer_
    int32_t r = glob( d.r )[ dev ];
    // ring-buffer cells are read LB lags at a time (independent loads), then consumed in
    // lag order and cleared
    for ( int l0 = ph.from; l0 < ph.to; l0 += LB )
    {
      double win_[ LB ];
  #pragma unroll
      for ( int q = 0; q < LB; ++q )
        if ( l0 + q < ph.to )
          win_[ q ] = C.get( l0 + q, 0, local );
  #pragma unroll
      for ( int q = 0; q < LB; ++q )
        if ( l0 + q < ph.to )
        {
          const int lag = l0 + q;
          C.put( lag, 0, local, 0.0 );
          if ( r == 0 )
          {
            y3 = P30 * y0I + P33 * y3 + win_[ q ];
            if ( refr_input and refr_buf != 0.0 )
            {
              y3 += refr_buf;
              refr_buf = 0.0;
            }
            y3 = ( y3 < V_min ? V_min : y3 );
          }
          else
          {
            if ( refr_input )
              refr_buf += win_[ q ] * ng_exp( -static_cast< double >( r ) * h / tau_m );
            --r;
          }
          if ( y3 >= V_th )
          {
            r = ps.refractory_counts;
            y3 = V_reset;
            bits[ lag >> 6 ] |= 1ull << ( lag & 63 );
          }
          if ( lag == next_sample )
          {
            store_sample( d, mcol, ph.origin + lag + 1, y3, 0.0, 0.0, 0.0 );
            next_sample += d.mm_int;
          }
        }
    }
    glob( d.v[ 0 ] )[ dev ] = y3;
    if ( refr_input )
      glob( d.v[ 1 ] )[ dev ] = refr_buf;
    glob( d.r )[ dev ] = r;
  }
};

template <>
struct DeviceNode< M_IAF_PSC_EXP >
{
  static constexpr bool WIDE = false;
  // iaf_psc_exp::handle( SpikeEvent& ) (iaf_psc_exp.cpp:351-362): positive weights to the
  // excitatory channel (1), others to the inhibitory one (0), signed
  __device__ __forceinline__ static void
  receive( double w, uint32_t& ch, double& v )
  {
    ch = w > 0 ? 1u : 0u;
    v = w;
  }
  template < bool WIN >
  __device__ __forceinline__ static void
  update( const NetDev& d, gref< const ParamSet > ps, uint32_t dev, uint32_t local, const Cells< WIN >& C,
    const SlicePhase& ph, uint64_t* bits, uint32_t mcol, int next_sample )
  {
    // models/iaf_psc_exp.cpp:288-349 (delta = 0: deterministic threshold)
    const double P22 = ps.p[ 0 ], P21ex = ps.p[ 1 ], P21in = ps.p[ 2 ], P20 = ps.p[ 3 ], P11ex = ps.p[ 4 ],
                 P11in = ps.p[ 5 ], IeI0 = ps.p[ 6 ], Theta = ps.p[ 7 ], V_reset = ps.p[ 8 ], exI1 = ps.p[ 9 ];
    double V = glob( d.v[ 0 ] )[ dev ];
    double iex = glob( d.v[ 1 ] )[ dev ];
    double iin = glob( d.v[ 2 ] )[ dev ];
    int32_t r = glob( d.r )[ dev ];
    for ( int l0 = ph.from; l0 < ph.to; l0 += LB )
    {
      double w_in[ LB ], w_ex[ LB ];
  #pragma unroll
      for ( int q = 0; q < LB; ++q )
        if ( l0 + q < ph.to )
        {
          w_in[ q ] = C.get( l0 + q, 0, local );
          w_ex[ q ] = C.get( l0 + q, 1, local );
        }
  #pragma unroll
      for ( int q = 0; q < LB; ++q )
        if ( l0 + q < ph.to )
        {
          const int lag = l0 + q;
          if ( r == 0 )
            V = V * P22 + iex * P21ex + iin * P21in + IeI0 * P20;
          else
            --r;
          iex *= P11ex;
          iin *= P11in;
          iex += exI1;
          iex += w_ex[ q ];
          iin += w_in[ q ];
          if ( V >= Theta )
          {
            r = ps.refractory_counts;
            V = V_reset;
            bits[ lag >> 6 ] |= 1ull << ( lag & 63 );
          }
          C.put( lag, 0, local, 0.0 );
          C.put( lag, 1, local, 0.0 );
          if ( lag == next_sample )
          {
            store_sample( d, mcol, ph.origin + lag + 1, V, iex, iin, 0.0 );
            next_sample += d.mm_int;
          }
        }
    }
    glob( d.v[ 0 ] )[ dev ] = V;
    glob( d.v[ 1 ] )[ dev ] = iex;
    glob( d.v[ 2 ] )[ dev ] = iin;
    glob( d.r )[ dev ] = r;
  }
};

template <>
struct DeviceNode< M_AEIF_COND_EXP >
{
  // The RKF45 stepper's state and coefficients want about 220 VGPRs. Held to 128 it spills
  // about 95 of them to scratch, and two workgroups per CU still run the batched sweep (half
  // its instances aeif_cond_exp) 37 % faster than one workgroup without spills.
  static constexpr bool WIDE = false;
  // aeif_cond_exp::handle( SpikeEvent& ) (aeif_cond_exp.cpp:526-541): positive weights to the
  // excitatory conductance (0), others negated to the inhibitory one (1)
  __device__ __forceinline__ static void
  receive( double w, uint32_t& ch, double& v )
  {
    if ( w > 0.0 )
    {
      ch = 0;
      v = w;
    }
    else
    {
      ch = 1;
      v = -w;
    }
  }
  template < bool WIN >
  __device__ __forceinline__ static void
  update( const NetDev& d, gref< const ParamSet > ps, uint32_t dev, uint32_t local, const Cells< WIN >& C,
    const SlicePhase& ph, uint64_t* bits, uint32_t mcol, int next_sample )
  {
    // models/aeif_cond_exp.cpp:443-523
    AeifPar A;
    A.V_peak = ps.p[ 0 ];
    A.V_reset = ps.p[ 1 ];
    A.g_L = ps.p[ 2 ];
    A.C_m = ps.p[ 3 ];
    A.E_ex = ps.p[ 4 ];
    A.E_in = ps.p[ 5 ];
    A.E_L = ps.p[ 6 ];
    A.Delta_T = ps.p[ 7 ];
    A.tau_w = ps.p[ 8 ];
    A.a = ps.p[ 9 ];
    A.b = ps.p[ 10 ];
    A.V_th = ps.p[ 11 ];
    A.tau_ex = ps.p[ 12 ];
    A.tau_in = ps.p[ 13 ];
    A.I_e = ps.p[ 14 ];
    A.tol = ps.p[ 15 ];
    const double V_peak_eff = ps.p[ 16 ];
    const double hstep = ps.p[ 17 ];
    double y[ 4 ] = { glob( d.v[ 0 ] )[ dev ], glob( d.v[ 1 ] )[ dev ], glob( d.v[ 2 ] )[ dev ],
      glob( d.v[ 3 ] )[ dev ] };
    double hint = glob( d.v[ 4 ] )[ dev ];
    int32_t r = glob( d.r )[ dev ];
    // One loop over (lag, accepted step): each iteration makes one adaptive step of the
    // lane's neuron, or, once the neuron has reached the end of its lag, closes the lag.
    // Lanes advance through their own lags independently, so a wave costs its slowest
    // neuron's total number of steps, not the sum over lags of the slowest per lag. Each
    // neuron's sequence of operations is that of the lag loop with an inner step loop.
    // More than AEIF_MAX_STEPS accepted steps within one time step stop the lane's lag and
    // raise ERR_AEIF_STEPS, which fails the run (the oracle stops with the same limit).
    int lag = ph.from;
    double tt = 0.0;
    int guard = 0;
    while ( lag < ph.to )
    {
      if ( tt < hstep and guard >= AEIF_MAX_STEPS )
      {
        g_or( &d.counters[ 0 ], ERR_AEIF_STEPS );
        tt = hstep;
      }
      if ( tt < hstep )
      {
        ++guard;
        aeif_evolve( A, r > 0, &tt, hstep, &hint, y );
        if ( y[ 0 ] < -1e3 or y[ 3 ] < -1e6 or y[ 3 ] > 1e6 )
        {
          g_or( &d.counters[ 0 ], ERR_AEIF_INSTABILITY );
          tt = hstep;
        }
        else if ( r > 0 )
          y[ 0 ] = A.V_reset;
        else if ( y[ 0 ] >= V_peak_eff )
        {
          y[ 0 ] = A.V_reset;
          y[ 3 ] += A.b;
          r = ps.refractory_counts > 0 ? ps.refractory_counts + 1 : 0;
          const uint64_t bit = 1ull << ( lag & 63 );
          uint64_t had = 0;
  #pragma unroll
          for ( int w = 0; w < 4; ++w )  // lag differs per lane: no dynamic register index
            if ( ( lag >> 6 ) == w )
            {
              had = bits[ w ] & bit;
              bits[ w ] |= bit;
            }
          if ( had )
            g_or( &d.counters[ 0 ], ERR_AEIF_DOUBLE_SPIKE );
        }
      }
      else
      {
        if ( r > 0 )
          --r;
        y[ 1 ] += C.get( lag, 0, local );
        y[ 2 ] += C.get( lag, 1, local );
        C.put( lag, 0, local, 0.0 );
        C.put( lag, 1, local, 0.0 );
        if ( lag == next_sample )
        {
          store_sample( d, mcol, ph.origin + lag + 1, y[ 0 ], y[ 1 ], y[ 2 ], y[ 3 ] );
          next_sample += d.mm_int;
        }
        ++lag;
        tt = 0.0;
        guard = 0;
      }
    }
    for ( int q = 0; q < 4; ++q )
      glob( d.v[ q ] )[ dev ] = y[ q ];
    glob( d.v[ 4 ] )[ dev ] = hint;
    glob( d.r )[ dev ] = r;
  }
};

template <>
struct DeviceNode< M_IAF_PSC_ALPHA >
{
  static constexpr bool WIDE = false;
  // iaf_psc_alpha::handle( SpikeEvent& ) (iaf_psc_alpha.cpp:373-383): positive weights to
  // the excitatory channel (1), others to the inhibitory one (0), signed
  __device__ __forceinline__ static void
  receive( double w, uint32_t& ch, double& v )
  {
    ch = w > 0 ? 1u : 0u;
    v = w;
  }
  template < bool WIN >
  __device__ __forceinline__ static void
  update( const NetDev& d, gref< const ParamSet > ps, uint32_t dev, uint32_t local, const Cells< WIN >& C,
    const SlicePhase& ph, uint64_t* bits, uint32_t mcol, int next_sample )
  {
    // models/iaf_psc_alpha.cpp:305-370 (the current input y0 is zero: no current sources)
    const double P30 = ps.p[ 0 ], y0I = ps.p[ 1 ], P31ex = ps.p[ 2 ], P32ex = ps.p[ 3 ], P31in = ps.p[ 4 ],
                 P32in = ps.p[ 5 ], expm1_tau_m = ps.p[ 6 ], P21ex = ps.p[ 7 ], P11ex = ps.p[ 8 ],
                 P21in = ps.p[ 9 ], P11in = ps.p[ 10 ], EPSCInit = ps.p[ 11 ], IPSCInit = ps.p[ 12 ],
                 Theta = ps.p[ 13 ], V_reset = ps.p[ 14 ], LowerBound = ps.p[ 15 ];
    double V = glob( d.v[ 0 ] )[ dev ];
    double iex = glob( d.v[ 1 ] )[ dev ];
    double iin = glob( d.v[ 2 ] )[ dev ];
    double diex = glob( d.v[ 3 ] )[ dev ];
    double diin = glob( d.v[ 4 ] )[ dev ];
    int32_t r = glob( d.r )[ dev ];
    // ring-buffer cells read LA lags at a time (its 16 propagators and 5 state variables
    // leave fewer registers for them than the other iaf models)
    constexpr int LA = 4;
    for ( int l0 = ph.from; l0 < ph.to; l0 += LA )
    {
      double w_in[ LA ], w_ex[ LA ];
#pragma unroll
      for ( int q = 0; q < LA; ++q )
        if ( l0 + q < ph.to )
        {
          w_in[ q ] = C.get( l0 + q, 0, local );
          w_ex[ q ] = C.get( l0 + q, 1, local );
        }
#pragma unroll
      for ( int q = 0; q < LA; ++q )
        if ( l0 + q < ph.to )
        {
          const int lag = l0 + q;
          if ( r == 0 )
          {
            V = P30 * y0I + P31ex * diex + P32ex * iex + P31in * diin + P32in * iin + expm1_tau_m * V + V;
            V = ( V < LowerBound ? LowerBound : V );
          }
          else
            --r;
          iex = P21ex * diex + P11ex * iex;  // P22_ex = P11_ex
          diex *= P11ex;
          diex += EPSCInit * w_ex[ q ];
          iin = P21in * diin + P11in * iin;
          diin *= P11in;
          diin += IPSCInit * w_in[ q ];
          if ( V >= Theta )
          {
            r = ps.refractory_counts;
            V = V_reset;
            bits[ lag >> 6 ] |= 1ull << ( lag & 63 );
          }
          C.put( lag, 0, local, 0.0 );
          C.put( lag, 1, local, 0.0 );
          if ( lag == next_sample )
          {
            store_sample( d, mcol, ph.origin + lag + 1, V, iex, iin, 0.0 );
            next_sample += d.mm_int;
          }
        }
    }
    glob( d.v[ 0 ] )[ dev ] = V;
    glob( d.v[ 1 ] )[ dev ] = iex;
    glob( d.v[ 2 ] )[ dev ] = iin;
    glob( d.v[ 3 ] )[ dev ] = diex;
    glob( d.v[ 4 ] )[ dev ] = diin;
    glob( d.r )[ dev ] = r;
  }
};

// models whose update needs the wide register budget
#define NG_WIDE( ID ) | ( DeviceNode< ( ID ) >::WIDE ? ( 1 << ( ID ) ) : 0 )
constexpr int WIDE_MODELS = 0 NG_DEVICE_MODELS( NG_WIDE );
#undef NG_WIDE

// handle( SpikeEvent& ) of model m
__device__ __forceinline__ void
node_receive( int m, double w, uint32_t& ch, double& v )
{
  ch = 0;
  v = w;
#define NG_RECEIVE( ID )                           \
  if ( m == ( ID ) )                               \
  {                                                \
    DeviceNode< ( ID ) >::receive( w, ch, v );     \
    return;                                        \
  }
  NG_DEVICE_MODELS( NG_RECEIVE )
#undef NG_RECEIVE
}

}  // namespace nestgpu
