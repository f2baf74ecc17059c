// ng_model.h -- neuron models as plug-ins: the host half of nest::Node on this path.
//
// NEST adds a neuron model by deriving from nest::Node (nestkernel/node.h: pre_run_hook :242,
// update( origin, from, to ) :297, set_status / get_status :337-346, handles_test_event :397,
// handle( SpikeEvent& ) :572) and registering the class under a name with
// kernel().model_manager.register_node_model< T >( name ) (nestkernel/nest_impl.h:40-45,
// model_manager_impl.h:42-55). On this path the update of every neuron runs inside the fused
// slice kernel, so a model has two halves:
//   - host (a NodeModel, this header): the model's name, its parameter and state defaults
//     (Parameters_(), State_()), set_status / get_status over its keys, the state variables it
//     records, and pre_run_hook, which turns the parameters into the ParamSet record that the
//     device update reads;
//   - device (ng_models_dev.h): update over the lags [from, to) of a slice and the receptor
//     channel a spike event of a given weight goes to (handle( SpikeEvent& )).
// Create, SetStatus, GetStatus, the multimeter and the store build reach a model only through
// ModelManager; a model registered with ModelManager::register_node_model and given a device
// half is simulated like the built-in ones.
#pragma once

#include <cstdint>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "ng_device.h"
#include "ng_json.h"

namespace nestgpu
{

// parameters of one neuron: the fields of the built-in models, and free slots for others
struct NeuronParams
{
  int32_t model;
  double tau_m, c_m, t_ref, E_L, I_e, V_th, V_min, V_reset;  // iaf: V values relative to E_L
  double tau_ex, tau_in;
  double V_peak, g_L, E_ex, E_in, Delta_T, tau_w, a, b, tol;  // aeif: absolute
  double delta, rho;
  double tau_minus;  // ArchivingNode: trace time constant of incoming stdp synapses
  int32_t refr_input;
  double extra[ 16 ];  // parameters of registered models without a named field
};

// state of one neuron: the device keeps fields v0..v4 and r per neuron
struct NeuronState
{
  double v0, v1, v2, v3, v4;  // delta: y3 | exp: V, i_ex, i_in | aeif: V, g_ex, g_in, w, h_int
  int32_t r;
};

class NodeModel
{
public:
  NodeModel( std::string name, int id, int channels )
    : name_( std::move( name ) )
    , id_( id )
    , channels_( channels )
  {
  }
  virtual ~NodeModel() = default;

  const std::string&
  name() const
  {
    return name_;
  }
  // the model id the device update dispatches on (ParamSet::model, NetDev::model)
  int
  id() const
  {
    return id_;
  }
  // ring-buffer channels of a target (receptor channels of handle( SpikeEvent& ))
  int
  channels() const
  {
    return channels_;
  }

  // Parameters_() and State_( const Parameters_& )
  virtual void defaults( NeuronParams& p, NeuronState& s ) const = 0;
  // Parameters_::set then State_::set with the checks the model makes, in its order: every
  // key the model reads is added to `used`; nothing changes unless every check passes
  virtual void set_status( NeuronParams& p, NeuronState& s, const Json& d, std::set< std::string >& used ) const = 0;
  // a parameter by its status key; false when the model has no such parameter
  virtual bool get_param( const NeuronParams& p, const std::string& key, double& v ) const = 0;
  // the device state field (v0..v3) of a state variable, -1 when the model has none of that
  // name; the variable's value is the field's plus state_offset
  virtual int state_field( const std::string& key ) const = 0;
  virtual double
  state_offset( const NeuronParams&, int ) const
  {
    return 0.0;
  }
  // the variables a multimeter may record (recordablesMap)
  virtual std::vector< std::string > recordables() const = 0;
  // pre_run_hook for resolution h (ms): the model's part of the device parameter set
  // (ps.p[], zeroed before); refractory_counts and the drive are filled by the caller
  virtual void pre_run_hook( const NeuronParams& p, double h, ParamSet& ps ) const = 0;
  // the device state fields v0..v4 at the start of the run
  virtual void
  device_state( const NeuronParams&, const NeuronState& s, double, double out[ 5 ] ) const
  {
    out[ 0 ] = s.v0;
    out[ 1 ] = s.v1;
    out[ 2 ] = s.v2;
    out[ 3 ] = s.v3;
    out[ 4 ] = s.v4;
  }

private:
  std::string name_;
  int id_;
  int channels_;
};

// the neuron model registry (ModelManager's node models)
class ModelManager
{
public:
  static ModelManager& instance();

  // registers a NodeModel under its name; the name must be new and its id the next device id
  template < class T, class... Args >
  const NodeModel&
  register_node_model( Args&&... args )
  {
    return add( std::make_unique< T >( std::forward< Args >( args )... ) );
  }
  // the model of that name, or nullptr
  const NodeModel* find( const std::string& name ) const;
  // the model of a device id
  const NodeModel& get( int id ) const;
  std::vector< std::string > names() const;
  int
  size() const
  {
    return static_cast< int >( models_.size() );
  }

private:
  ModelManager();
  const NodeModel& add( std::unique_ptr< NodeModel > m );
  std::vector< std::unique_ptr< NodeModel > > models_;
};

}  // namespace nestgpu
