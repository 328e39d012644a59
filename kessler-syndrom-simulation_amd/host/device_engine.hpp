// RAII owner of one engine context (include/kessler_hip.h), shared by a
// Simulation and the Screener attached to it. Errors become
// std::runtime_error on this side of the C ABI.

#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kessler_hip.h"
#include "collision.hpp"

namespace kessler {

class DeviceEngine {
public:
    // nullptr and the reason in `message` when no device can be used.
    static std::shared_ptr<DeviceEngine> create(int device, std::string* message);
    ~DeviceEngine();
    DeviceEngine(const DeviceEngine&) = delete;
    DeviceEngine& operator=(const DeviceEngine&) = delete;

    khip_ctx* ctx() const { return ctx_; }
    const std::string& name() const { return name_; }
    void check(int rc, const char* what) const;

    // Screening requested for the next step (Screener::begin) and the
    // configuration it uses.
    bool screen_next = false;
    ScreeningConfig cfg;
    // Results of the last screened step, taken by Screener::finish.
    bool have_results = false;
    std::vector<khip_event> events;
    khip_stats stats{};

private:
    DeviceEngine() = default;
    khip_ctx* ctx_ = nullptr;
    std::string name_;
};

}  // namespace kessler
