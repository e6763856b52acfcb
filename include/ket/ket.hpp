// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// Umbrella header of the ket API served by the MI355X backend.
#pragma once

#include <ket/backends/backend.hpp>
#include <ket/circuit.hpp>
#include <ket/dag.hpp>
#include <ket/gate.hpp>
#include <ket/measurement.hpp>
#include <ket/observable.hpp>
#include <ket/parallelism.hpp>
#include <ket/qasm.hpp>
#include <ket/qubit.hpp>
#include <ket/simulator.hpp>
#include <ket/state.hpp>
#include <ket/stepper.hpp>
#include <ket/version.hpp>
#include <ket/amd.hpp>
