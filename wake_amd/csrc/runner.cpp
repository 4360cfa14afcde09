// runner.cpp -- SimulationRunner (the reference's runner.cpp:111-184 loop).
#include "wake/runner/simulation_runner.hpp"

#include <chrono>
#include <filesystem>
#include <iostream>
#include <stdexcept>

namespace wake {

SimulationRunner::SimulationRunner(std::shared_ptr<Scenario> scenario, const RunnerConfig& config)
    : scenario_(std::move(scenario)), config_(config) {}

SimulationRunner::~SimulationRunner() {
  stop();
  if (worker_thread_.joinable()) worker_thread_.join();
}

void SimulationRunner::run() {
  if (!scenario_) throw std::runtime_error("Cannot run null scenario in SimulationRunner");
  is_running_ = true;
  should_stop_ = false;
  std::filesystem::create_directories(config_.export_dir);
  scenario_->setup();
  double total_mlups = 0.0;
  uint64_t measured = 0;
  int frame = 0;
  while (!should_stop_ && !scenario_->is_finished() && scenario_->current_step() < config_.max_steps) {
    while (is_paused_ && !should_stop_) std::this_thread::sleep_for(std::chrono::milliseconds(10));
    if (should_stop_) break;
    const auto t0 = std::chrono::steady_clock::now();
    scenario_->step();
    const uint64_t step = scenario_->current_step();
    step_count_ = step;
    const double mlups = scenario_->current_mlups();
    if (mlups > 0.0) {
      total_mlups += mlups;
      ++measured;
      avg_mlups_ = total_mlups / static_cast<double>(measured);
    }
    if (step_callback_) step_callback_(step, mlups);
    if (config_.live_terminal_render && config_.render_interval && step % config_.render_interval == 0)
      scenario_->render_terminal(std::cout);
    if (config_.export_interval && step % config_.export_interval == 0)
      scenario_->export_frame(config_.export_dir, frame++);
    if (config_.target_fps > 0.0) {
      const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      const double budget = 1.0 / config_.target_fps;
      if (elapsed < budget) std::this_thread::sleep_for(std::chrono::duration<double>(budget - elapsed));
    }
  }
  scenario_->export_frame(config_.export_dir, frame);
  scenario_->export_telemetry(config_.export_dir);
  is_running_ = false;
}

void SimulationRunner::start_async() {
  if (is_running_) return;
  is_running_ = true;
  worker_thread_ = std::thread([this]() { run(); });
}

void SimulationRunner::pause() { is_paused_ = true; }
void SimulationRunner::resume() { is_paused_ = false; }
void SimulationRunner::stop() {
  should_stop_ = true;
  is_paused_ = false;
}
void SimulationRunner::wait() {
  if (worker_thread_.joinable()) worker_thread_.join();
}
uint64_t SimulationRunner::current_step() const { return scenario_ ? scenario_->current_step() : step_count_.load(); }

}  // namespace wake
