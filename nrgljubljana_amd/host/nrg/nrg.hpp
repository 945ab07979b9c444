// nrg.hpp - the host layer: reference-shaped C++ interfaces over the nrgcore C ABI.
#pragma once
#include "traits.hpp"
#include "device.hpp"
#include "eigen.hpp"
#include "diag.hpp"
#include "hamiltonian.hpp"
#include "numerics.hpp"
#include "recalc.hpp"
#include "measurements.hpp"
#include "dmnrg.hpp"
#include "spectrum.hpp"
#include "store.hpp"
#include "read_input.hpp"
