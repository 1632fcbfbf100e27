// Device side of the smoothed-aggregation setup (amg_device.hip): the sparse products of a level
// with the aggregation done on the host.
#pragma once

#include "common.hpp"
#include "hip_util.hpp"

#include <memory>
#include <vector>

namespace cfdg {

// Holds the current level's operator on the device (level 0 at construction). level() forms
// the smoothed prolongation P (pruned), R = P^T and the Galerkin operator R A P (pruned) from the
// level's inverse diagonal, aggregates and damping, returns the three on the host (bit-identical
// to the host products) and keeps the coarse operator on the device as the next level's.
class DeviceSetup {
public:
    explicit DeviceSetup(const Csr& a0);
    ~DeviceSetup();
    DeviceSetup(const DeviceSetup&) = delete;
    DeviceSetup& operator=(const DeviceSetup&) = delete;

    void level(const std::vector<double>& inv_diag, const std::vector<int>& agg, int n_agg, double omega, Csr& p,
               Csr& r, Csr& coarse);
    // Passes 1 and 2 of the greedy aggregation (amg.cpp:12-63) of the current level on the device,
    // bit-identical to the sequential loops: the aggregate of every node, -1 for the nodes pass 3
    // must place; n_roots = the aggregates of pass 1.
    std::vector<int> aggregate12(const std::vector<double>& d, double theta, int& n_roots);

    struct DCsr;

private:
    Stream stream_;
    std::unique_ptr<DCsr> cur_;
};

} // namespace cfdg
