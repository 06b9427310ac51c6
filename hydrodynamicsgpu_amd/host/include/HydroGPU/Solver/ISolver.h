// HydroGPU/Solver/ISolver.h -- (include/HydroGPU/Solver/ISolver.h:12-17)
#pragma once
#include <memory>
#include <string>

namespace HydroGPU {
namespace Equation {
struct Equation;
}
namespace Solver {

struct ISolver {
  virtual ~ISolver() {}
  virtual void init() = 0;
  virtual void resetState() = 0;
  virtual std::string name() const = 0;
  virtual std::shared_ptr<Equation::Equation> getEquation() const = 0;
};

}  // namespace Solver
}  // namespace HydroGPU
