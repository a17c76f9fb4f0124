// String-keyed singleton factories.  Same templates, names and semantics as
// the reference's include/factory.cuh:8-89 (Singleton, Factory,
// createObject, registerCreationFunction); an unknown id throws
// gvm::FactoryError instead of exiting, so callers decide how to fail.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <typeinfo>

namespace gvm {
class FactoryError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};
}  // namespace gvm

template <class T>
class Singleton {
 public:
  static T& Instance() {
    static T instance;
    return instance;
  }

 private:
  Singleton() {}
  Singleton(T const&) = delete;
  void operator=(T const&) = delete;
};

template <class AbstractProduct, class IdentifierType,
          class ProductCreator = AbstractProduct* (*)()>
class Factory {
 public:
  bool Register(const IdentifierType& id, ProductCreator creator) {
    return associations_.insert(typename AssocMap::value_type(id, creator)).second;
  }
  bool Unregister(const IdentifierType& id) { return associations_.erase(id) == 1; }
  AbstractProduct* CreateObject(const IdentifierType& id) {
    auto it = associations_.find(id);
    if (it != associations_.end()) return (it->second)();
    throw gvm::FactoryError(std::string("Unknown object type passed to Factory of class ") +
                            typeid(AbstractProduct).name());
  }

 private:
  typedef std::map<IdentifierType, ProductCreator> AssocMap;
  AssocMap associations_;
};

template <class T, class V>
T* createObject(V value) {
  return Singleton<Factory<T, V>>::Instance().CreateObject(value);
}

template <class T, class V, class Creator = T* (*)()>
bool registerCreationFunction(V value, Creator function) {
  return Singleton<Factory<T, V>>::Instance().Register(value, function);
}
