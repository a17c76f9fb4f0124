// Optimiser interface: the reference's include/classes/optimizer.cuh:5-60.
// Registered implementations: "CG-FRPRMN" (ConjugateGradient) and
// "CG-LBFGS" (LBFGS).  Tolerances are kept as float like the reference; the
// line-search and direction scalars are FP64.
#pragma once
#include <chrono>
#include <vector>

#include "gvm/image.h"
#include "gvm/objectivefunction.h"

class Optimizer {
 public:
  virtual ~Optimizer() {}
  virtual void allocateMemoryGpu() = 0;
  virtual void deallocateMemoryGpu() = 0;
  virtual void optimize() = 0;
  virtual int getK() { return 0; }
  virtual void setK(int K) {}

  Optimizer() : ftol(1E-12f), gtol(1E-12f), total_iterations(500) {}
  Optimizer(int total_iterations, float ftol)
      : ftol(ftol), gtol(1E-12f), total_iterations(total_iterations) {}
  Optimizer(int total_iterations, float ftol, float gtol)
      : ftol(ftol), gtol(gtol), total_iterations(total_iterations) {}

  float getFtol() { return this->ftol; }
  float getGtol() { return this->gtol; }
  int getCurrentIteration() { return this->current_iteration; }
  void setImage(Image* image) { this->image = image; }
  void setObjectiveFunction(ObjectiveFunction* of) { this->of = of; }
  void setFlag(int flag) { this->flag = flag; }
  void setFTol(float ftol) { this->ftol = ftol; }
  void setGTol(float gtol) { this->gtol = gtol; }
  void setTotalIterations(int iterations) { this->total_iterations = iterations; }
  ObjectiveFunction* getObjectiveFunction() { return this->of; }

  // Trajectory record (not in the reference): function value after each
  // line search, and why the loop ended (0 iteration limit, 1 ftol,
  // 2 gradient tolerance, 3 gg == 0).
  const std::vector<double>& getTrajectory() const { return trajectory; }
  // Wall-clock seconds from the start of optimize() to each trajectory entry
  // (the value is on the host then, so the device work is complete).
  const std::vector<double>& getTrajectorySeconds() const { return trajectory_seconds; }
  int getExitReason() const { return exit_reason; }
  double getFinalValue() const { return final_value; }

 protected:
  ObjectiveFunction* of = nullptr;
  Image* image = nullptr;
  int flag = 0;
  float ftol;
  float gtol;
  int total_iterations;
  int current_iteration = 0;
  std::vector<double> trajectory;
  std::vector<double> trajectory_seconds;
  std::chrono::steady_clock::time_point start_time;
  void startTrajectory() {
    trajectory.clear();
    trajectory_seconds.clear();
    start_time = std::chrono::steady_clock::now();
  }
  void recordTrajectory(double value) {
    trajectory.push_back(value);
    trajectory_seconds.push_back(
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start_time).count());
  }
  int exit_reason = 0;
  double final_value = 0.0;
};
