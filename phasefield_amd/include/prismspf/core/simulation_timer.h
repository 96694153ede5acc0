// Simulated-time bookkeeping handed to the equation hooks
// (reference: include/prismspf/core/simulation_timer.h).  It is a plain struct so a
// copy travels to the device as a kernel argument.
#pragma once

#include <prismspf/config.h>

PRISMS_PF_BEGIN_NAMESPACE

class SimulationTimer
{
public:
  SimulationTimer() = default;

  PPF_HD explicit SimulationTimer(double step_size, double start_time = 0.0)
    : time_step_size(step_size)
    , current_time(start_time)
  {}

  [[nodiscard]] PPF_HD unsigned int
  get_increment() const
  {
    return current_increment;
  }

  [[nodiscard]] PPF_HD double
  get_time() const
  {
    return current_time;
  }

  [[nodiscard]] PPF_HD double
  get_timestep() const
  {
    return time_step_size;
  }

  void
  increment(double step_size)
  {
    current_increment++;
    current_time += step_size;
  }

  void
  increment()
  {
    increment(time_step_size);
  }

  void
  set_timestep(double step_size)
  {
    time_step_size = step_size;
  }

  void
  set_time(double time)
  {
    current_time = time;
  }

  void
  set_increment(unsigned int increment)
  {
    current_increment = increment;
  }

  void
  reset()
  {
    current_increment = 0;
    current_time      = 0.0;
  }

private:
  unsigned int current_increment = 0;
  double       current_time      = 0.0;
  double       time_step_size    = 0.0;
};

PRISMS_PF_END_NAMESPACE
