// dacti/driver/observer/common.hpp -- observer base (driver/observer/
// common.hpp:8-40 of the reference): a name, integral variables (one value
// per observation, written to the observables CSV) and field variables (one
// value per active cell, written to the frames).
#ifndef DACTI_OBSERVER_COMMON_HPP
#define DACTI_OBSERVER_COMMON_HPP

#include <string>
#include <utility>
#include <vector>

#include "../../core/types.hpp"

namespace dacti::observer {

class Observer {
protected:
    std::string _observer_name;
    std::vector<std::string> _IntegVar_names, _FieldVar_names;
    size_t _NumIntegVars, _NumFieldVars;
    mutable std::vector<scalar_t> _IntegVars;
    mutable std::vector<std::vector<scalar_t>> _FieldVars;

public:
    explicit Observer(std::string observer_name = "Default Observer", std::vector<std::string> IntegVar_names = {},
                      std::vector<std::string> FieldVar_names = {})
        : _observer_name(std::move(observer_name)),
          _IntegVar_names(std::move(IntegVar_names)),
          _FieldVar_names(std::move(FieldVar_names)),
          _NumIntegVars(_IntegVar_names.size()),
          _NumFieldVars(_FieldVar_names.size()),
          _IntegVars(_NumIntegVars),
          _FieldVars(_NumFieldVars) {}

    size_t NumIntegVars() const { return _NumIntegVars; }
    size_t NumFieldVars() const { return _NumFieldVars; }
    std::string name() const { return _observer_name; }
    std::vector<std::string> IntegVar_names() const { return _IntegVar_names; }
    std::vector<std::string> FieldVar_names() const { return _FieldVar_names; }
    std::vector<scalar_t> IntegVars() const { return _IntegVars; }
    std::vector<std::vector<scalar_t>> FieldVars() const { return _FieldVars; }
};

}  // namespace dacti::observer

#endif
