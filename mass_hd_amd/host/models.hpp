// models.hpp -- the Model plugin interface of the host program.
//
// Same interface as the reference (src/models.hpp:30-216): a Model is
// constructed by a registered factory from (LX, LY, BC), receives its options
// through GetOptions(), then the driver calls Initialize(), Configure(),
// Pre(), Step() in the time loop, RuntimeChecks()/RuntimeStats() every ninfo
// steps, and Post().  serialize_params/serialize_frame are templates in the
// concrete model, reached through ModelPtr's type-erased trampoline.
#pragma once

#include <functional>
#include <memory>
#include <string>

#include "options.hpp"
#include "serialization.hpp"

class Model {
 protected:
  unsigned LX = 0, LY = 0, BC = 0;
  unsigned DomainSize = 0;

 public:
  Model() = default;
  Model(unsigned LX_, unsigned LY_, unsigned BC_)
      : LX(LX_), LY(LY_), BC(BC_), DomainSize(LX_ * LY_) {}
  virtual ~Model() {}
  // derived quantities from the options; may throw
  virtual void Initialize() = 0;
  // initial configuration; may throw
  virtual void Configure() = 0;
  // model-specific options
  virtual option_list GetOptions() = 0;
  virtual void Pre() {}
  // one time step; does not throw
  virtual void Step() = 0;
  virtual void Post() {}
  // every ninfo steps; throws warning_msg / error_msg
  virtual void RuntimeChecks() {}
  // every ninfo steps for verbose > 1
  virtual void RuntimeStats() {}
  virtual void PreRunStats() {}
};

// Type-erased access to the templated serializers of the concrete model.
struct ModelPtr {
  struct Serializer {
    Model* ptr;
    explicit Serializer(Model* p) : ptr(p) {}
    virtual ~Serializer() {}
    virtual void serialize_params(oarchive&) = 0;
    virtual void serialize_frame(oarchive&) = 0;
    virtual void serialize_frame(frame_values&) = 0;
  };
  template <class Derived>
  struct SerializerT : Serializer {
    explicit SerializerT(Derived* p) : Serializer(p) {}
    void serialize_params(oarchive& ar) override { static_cast<Derived*>(ptr)->serialize_params(ar); }
    void serialize_frame(oarchive& ar) override { static_cast<Derived*>(ptr)->serialize_frame(ar); }
    void serialize_frame(frame_values& ar) override {
      static_cast<Derived*>(ptr)->serialize_frame(ar);
    }
  };

  std::shared_ptr<Serializer> ser;
  std::shared_ptr<Model> owner;

  ModelPtr() = default;
  template <class Derived>
  explicit ModelPtr(Derived* m) : ser(std::make_shared<SerializerT<Derived>>(m)), owner(m) {}

  Model* operator->() { return owner.get(); }
  const Model* operator->() const { return owner.get(); }
  explicit operator bool() const { return (bool)owner; }
};

using Factory = std::function<ModelPtr(unsigned, unsigned, unsigned)>;

void add_model(const std::string& name, const std::string& desc, Factory f);

template <class T>
void declare_model(const std::string& name, const std::string& desc) {
  add_model(name, desc, [](unsigned LX, unsigned LY, unsigned BC) {
    return ModelPtr(new T(LX, LY, BC));
  });
}

ModelPtr NewModel(const std::string& name, unsigned LX, unsigned LY, unsigned BC);
std::string ListModels();
void DeclareModels();
