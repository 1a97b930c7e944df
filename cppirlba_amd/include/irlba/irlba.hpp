#ifndef IRLBA_HIP_IRLBA_HPP
#define IRLBA_HIP_IRLBA_HPP

// Umbrella header (include/irlba/irlba.hpp of the reference): the irlba::
// API over the MI355X engine.  Link with cppirlba_amd/lib/libirlba_hip.so.

#include "Matrix/centered.hpp"
#include "Matrix/hip.hpp"
#include "Matrix/interface.hpp"
#include "Matrix/scaled.hpp"
#include "Options.hpp"
#include "compute.hpp"
#include "pca.hpp"
#include "types.hpp"

#endif
