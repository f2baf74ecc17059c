/*
 * nestgpu.h -- C ABI of the MI355X point-neuron simulation step.
 *
 * The reference exposes its simulation kernel to Python through a Cython binding
 * (pynest/nestkernel_api.pyx) over the C++ API in nestkernel/nest.h. This header is the
 * drop-in counterpart of that binding surface: the same calls (create, connect,
 * connect_arrays, simulate, prepare/run/cleanup, kernel status, node status,
 * GetConnections, spike_recorder events) with dictionaries passed as JSON text, plain
 * pointers and sizes, integer status returns (0 = success) and the error message of the
 * last failing call readable through nestgpu_last_error(). Node ids are NEST's 1-based
 * global ids. No exception crosses this boundary. One host thread per kernel handle;
 * several handles may share a device.
 */
#ifndef NESTGPU_H
#define NESTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nestgpu_kernel nestgpu_kernel;

/* nest::init_nest + ResetKernel (nestkernel/nest.cpp:54; pynest nestkernel_api.pyx llapi_reset_kernel).
 * Binds the kernel to a HIP device; device memory is allocated at the first Prepare. */
int nestgpu_kernel_new( int device, nestgpu_kernel** out );
void nestgpu_kernel_free( nestgpu_kernel* k );
/* Message of the last failing call on k (or of the last failing call without a kernel when k is NULL). */
const char* nestgpu_last_error( const nestgpu_kernel* k );
/* nest::reset_kernel (nestkernel/nest.cpp:101; pynest/nestkernel_api.pyx:330 llapi_reset_kernel). */
int nestgpu_reset_kernel( nestgpu_kernel* k );

/* nest::set_kernel_status / get_kernel_status (nestkernel/nest.cpp:149,157; nestkernel_api.pyx:547,568):
 * keys resolution, rng_seed, total_num_virtual_procs, local_num_threads, rng_type, and the
 * device layout keys targets_per_block, delivery_mode ("auto", "window", "global"),
 * slices_per_chunk, kernel_timing. */
int nestgpu_set_kernel_status( nestgpu_kernel* k, const char* json );
/* Writes the status dictionary as JSON into buf (NUL-terminated); *len receives the full length. */
int nestgpu_get_kernel_status( nestgpu_kernel* k, char* buf, size_t cap, size_t* len );

/* nest::create (nestkernel/nest.cpp:362): models iaf_psc_delta, iaf_psc_exp, aeif_cond_exp,
 * spike_generator, spike_recorder, multimeter, voltmeter. *first_id receives the id of the first
 * new node. */
int nestgpu_create( nestgpu_kernel* k, const char* model, int64_t n, const char* params_json, int64_t* first_id );
/* nest::set_nc_status (nestkernel/nest.cpp:210) with one dictionary for every node (NodeCollection.set(dict)). */
int nestgpu_set_status( nestgpu_kernel* k, const int64_t* ids, int64_t n, const char* params_json );
/* NodeCollection.set(key=array): one value per node. */
int nestgpu_set_values( nestgpu_kernel* k, const int64_t* ids, int64_t n, const char* key, const double* values );
/* nest::get_nc_status (nestkernel/nest.cpp:169), NodeCollection.get(key): one value per node
 * (state is read back from the device after Simulate). */
int nestgpu_get_values( nestgpu_kernel* k, const int64_t* ids, int64_t n, const char* key, double* out );

/* nest::connect (nestkernel/nest.cpp:424): conn_spec rules one_to_one, all_to_all,
 * fixed_indegree, pairwise_bernoulli; syn_spec static_synapse with weight/delay numbers or
 * parameter objects ({"normal": {"mean", "std"}}, {"uniform": {"min", "max"}},
 * {"uniform_int": {"max"}}, {"sum": [a, b]}, {"product": [a, b]}, {"max": [a, v]},
 * {"min": [a, v]}, {"redraw": [a, min, max]}). Random draws follow NEST's per-VP streams. */
int nestgpu_connect( nestgpu_kernel* k, const int64_t* pre, int64_t n_pre, const int64_t* post, int64_t n_post,
  const char* conn_spec_json, const char* syn_spec_json );
/* nest::connect_arrays (nestkernel/nest.cpp:454); weights/delays may be NULL (defaults).
 * Every element is checked first: a call that fails (IllegalConnection, BadDelay) adds no
 * connection. */
int nestgpu_connect_arrays( nestgpu_kernel* k, const int64_t* sources, const int64_t* targets, const double* weights,
  const double* delays_ms, int64_t n );
/* spike_generator input as dense multiplicities: counts[s * n + i] spikes of generator
 * first_id + i with stamp (first_step + s + 1) * resolution. Equivalent to spike_times and
 * spike_multiplicities of models/spike_generator.cpp:85-140 for on-grid times. */
int nestgpu_set_spike_counts( nestgpu_kernel* k, int64_t first_id, int64_t n, int64_t first_step, int64_t n_steps,
  const uint8_t* counts );

/* nest::get_connections (nestkernel/nest.cpp:473), GetConnections: per VP in creation order,
 * then spike_generator connections. */
int64_t nestgpu_num_connections( nestgpu_kernel* k );
int nestgpu_get_connections( nestgpu_kernel* k, int64_t* sources, int64_t* targets, double* weights,
  int64_t* delay_steps, int32_t* vps );
/* The synapse model of each connection, in nestgpu_get_connections' order: 0 static_synapse,
 * 1 stdp_synapse (the synapse_model / synapse_modelid a connection's status reports,
 * nestkernel/connector_model_impl.h:105-106, connection_id.cpp:65). */
int nestgpu_get_connection_models( nestgpu_kernel* k, int32_t* models );

/* nest::prepare / run / cleanup / simulate (nestkernel/nest.cpp:499-534). */
int nestgpu_prepare( nestgpu_kernel* k );
int nestgpu_run( nestgpu_kernel* k, double t_ms );
int nestgpu_cleanup( nestgpu_kernel* k );
int nestgpu_simulate( nestgpu_kernel* k, double t_ms );

/* spike_recorder events (models/spike_recorder.cpp:100-112), sorted by (time, sender).
 * Call with senders == NULL to get the count in *n. */
int nestgpu_recorder_events( nestgpu_kernel* k, int64_t recorder_id, int64_t* senders, double* times, int64_t cap,
  int64_t* n );

/* multimeter / voltmeter events (models/multimeter.cpp:190-216 over the UniversalDataLogger
 * samples, nestkernel/universal_data_logger_impl.h:343-375), sorted by (time, sender):
 * values[i * n_fields + j] is record_from[j] of event i, with n_fields the length of the
 * multimeter's record_from list (nestgpu_node_status gives it). Call with senders == NULL
 * to get the count in *n. */
int nestgpu_multimeter_events( nestgpu_kernel* k, int64_t multimeter_id, int64_t* senders, double* times,
  double* values, int64_t cap, int64_t* n );
/* Status dictionary of one node as JSON (nest::get_nc_status, nestkernel/nest.cpp:169):
 * model and global_id, and a multimeter's interval, offset and record_from. Same buffer
 * convention as nestgpu_get_kernel_status. */
int nestgpu_node_status( nestgpu_kernel* k, int64_t id, char* buf, size_t cap, size_t* len );

/* Batched form: advances several prepared kernels on one device together, one launch per
 * phase per min_delay slice for all of them (no counterpart in the reference, which runs
 * one kernel per process). t_ms[i] is the simulation time of kernel i. */
int nestgpu_batch_simulate( nestgpu_kernel** ks, int n, const double* t_ms );

/* Variant of nestgpu_batch_simulate that allocates nothing (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Every kernel must already be prepared
 * (nestgpu_prepare, or an earlier Simulate) with no change since; the call then makes no
 * hipMalloc, hipFree, host-pinned or graph allocation: the launch records, block prefix and
 * status words of the slice kernels live in a fixed device array of the library, one per
 * device, set up at the first call on that device, and the slices are launched directly on
 * the device's null stream. At most 64 kernels per call (larger n is an error). Results are
 * identical to nestgpu_batch_simulate. Calls that overlap on one device share that array and
 * race; calls on one device must not run concurrently. */
int nestgpu_batch_simulate_noalloc( nestgpu_kernel** ks, int n, const double* t_ms );

/* Layout autotuner of the batched form (no counterpart in the reference): groups the kernels
 * (none simulated yet) into classes by neuron models and neuron count (one per octave), times
 * probe runs of probe_ms on up to 4 instances of each class for every pair of a candidate
 * targets_per_block (the target tile of a workgroup) and a candidate cap on the ring-buffer
 * slots staged in LDS (window_caps, 0 = as many as fit; NULL or n_win 0: that default only),
 * and sets the cheapest pair on every instance of the class (each keeps its delivery_mode). A
 * probe's cost is its block-time, wall time x min(blocks, 2 x compute units): its share of a
 * batch that fills the device. Probe runs are discarded: every kernel returns to its state
 * before the first Prepare. chosen[i] and chosen_win[i] (either may be NULL) receive the
 * values picked for kernel i. */
int nestgpu_batch_autotune( nestgpu_kernel** ks, int n, double probe_ms, const int* candidates, int n_cand,
  const int* window_caps, int n_win, int* chosen, int* chosen_win );

/* Served form (counterpart of nest-server, pynest/nest/server/hl_api_server.py, which runs
 * one kernel per request and serialises calls): a request loop on one device. Admitted
 * networks stay resident in HBM across their Simulate calls; all running requests advance
 * together in the same slice launches, and new requests join at chunk boundaries as soon as
 * the HBM budget (bytes) has room. Finished requests release their device memory; their
 * spikes and final state stay readable through the kernel handle. */
typedef struct nestgpu_server nestgpu_server;
int nestgpu_server_new( int device, uint64_t hbm_budget_bytes, int slices_per_chunk, nestgpu_server** out );
void nestgpu_server_free( nestgpu_server* s );
/* k: built on the host (Create/Connect), not prepared or simulated, owned by the caller and
 * untouched by it until the request is done; t_ms: its n_calls Simulate durations. */
int nestgpu_server_submit( nestgpu_server* s, nestgpu_kernel* k, const double* t_ms, int n_calls, int64_t* id );
/* Waits up to timeout_s (negative: forever). state: 0 queued, 1 admitted, 2 running, 3 done,
 * -1 failed. times[4] (may be NULL): submit, admit, start of the first run, done, in seconds
 * since the server started. events: recurrent synaptic events of the request. */
int nestgpu_server_wait( nestgpu_server* s, int64_t id, double timeout_s, int* state, double* times,
  uint64_t* events );
/* Open request (a session, as nest-server clients use the kernel): like nestgpu_server_submit,
 * but the network stays resident in HBM after its calls are done, until
 * nestgpu_server_close. nestgpu_server_extend appends a Simulate call of t_ms, which joins the
 * running batch at the next chunk boundary and continues from the resident state;
 * nestgpu_server_wait_calls waits until n calls are done (the kernel's state and events are
 * readable between calls). */
int nestgpu_server_submit_open( nestgpu_server* s, nestgpu_kernel* k, const double* t_ms, int n_calls, int64_t* id );
int nestgpu_server_extend( nestgpu_server* s, int64_t id, double t_ms );
int nestgpu_server_close( nestgpu_server* s, int64_t id );
int nestgpu_server_wait_calls( nestgpu_server* s, int64_t id, int64_t n, double timeout_s, int* state,
  int64_t* calls_done );
/* Server statistics as JSON (same buffer convention as nestgpu_get_kernel_status). */
int nestgpu_server_stats( nestgpu_server* s, char* buf, size_t cap, size_t* len );

/* Seeded Poisson drive counts, the input the spike_generator option poisson_rate produces:
 * out[s * n + i] for streams stream0 + i and emission steps step0 + s, lambda per step. */
int nestgpu_poisson_counts( uint64_t seed, double lambda, uint64_t stream0, int64_t n, int64_t step0, int64_t n_steps,
  uint8_t* out );

/* Number of visible HIP devices (0 without a GPU). */
int nestgpu_device_count( int* n );

#ifdef __cplusplus
}
#endif

#endif /* NESTGPU_H */
