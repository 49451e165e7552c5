// The context behind an sfp_context handle (shared by capi.cpp and capi_test.cpp).
#pragma once
#include "../../include/sfp.h"
#include "../../include/sfp_profile.h"
#include "solver.hpp"

#include <memory>
#include <string>

struct sfp_context_s {
    int device = 0;
    sfp::Dims d;
    double spacing = 0;
    double origin[3] = {0, 0, 0};
    sfp::MagGeom mg;
    sfp::SolverConfig cfg;
    sfp::Workspace ws;
    sfp::System pressure, magnetic;
    // device-resident state
    sfp::DevBuf<unsigned char> liquid;  // unknowns of the last projection
    sfp::DevBuf<double> mu[3];          // magnetic face coefficients
    sfp::DevBuf<double> htot[3];        // last total field (for vertex pressures)
    bool pressure_ready = false;
    long long pressure_unknowns = 0;
    bool magnetic_warm = false;
    bool field_ready = false;
    // staging for the host-pointer entry points
    sfp::DevBuf<double> s_phi, s_cell;
    // velocity, fraction, solid velocity (or normal), jump (or collider
    // velocity), field (or normal), collider velocity
    sfp::DevBuf<double> s_face[6][3];
    sfp::DevBuf<double> s_vert[3];      // positions, normals, pressures
    // post-projection steps: velocity before Collider::Enforce, validity masks
    sfp::DevBuf<double> post_old[3];
    sfp::DevBuf<unsigned char> post_valid[2][3];
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::unique_ptr<sfp::UNet> net;  // learned preconditioner weights, when given
    std::string last_error;
};
