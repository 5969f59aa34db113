// Compile-time checks of the C++ layer's Traits -> library-code mapping
// (compiled with -fsyntax-only by tests/test_cpp_layer.py).
#include "hemelb_amd.hpp"

using namespace hemelb;

// The reference's CMake defaults: D3Q15, LBGK, simple bounce-back, Nash.
using Default = lb::TraitsCodes<Traits<>>;
static_assert(Default::lattice == 15 && Default::kernel == HLB_LBGK, "defaults");
static_assert(Default::wall == HLB_WALL_SBB, "default wall");
static_assert(Default::inlet == HLB_IOLET_NASH && Default::outlet == HLB_IOLET_NASH, "iolets");

// The vessel-tree build: D3Q19 LBGK, BFL walls, Ladd inlet, Nash outlets.
using Tree = lb::TraitsCodes<Traits<lb::D3Q19, lb::LBGK, lb::Normal, lb::BulkStreamer,
                                    lb::BouzidiFirdaousLallemand, lb::LaddIolet,
                                    lb::NashZerothOrderPressure>>;
static_assert(Tree::lattice == 19 && Tree::kernel == HLB_LBGK && Tree::wall == HLB_WALL_BFL,
              "tree build");
static_assert(Tree::inlet == HLB_IOLET_LADD && Tree::outlet == HLB_IOLET_NASH, "tree iolets");

// MRT picks the d'Humieres basis of the lattice.
using Mrt15 = Traits<lb::D3Q15, lb::DHumieresMRT>;
static_assert(std::is_same<Mrt15::Kernel, lb::MRT<lb::DHumieresD3Q15MRTBasis>>::value, "basis");
static_assert(lb::TraitsCodes<Mrt15>::kernel == HLB_MRT, "MRT code");
using Mrt19 = Traits<lb::D3Q19, lb::DHumieresMRT>;
static_assert(std::is_same<Mrt19::Kernel::LatticeType, lb::D3Q19>::value, "MRT lattice");

// Streamers assembled from links, as StreamerTypeFactory does.
template <class C>
using BflByHand = lb::StreamerTypeFactory<lb::BouzidiFirdaousLallemandLink<C>, lb::NullLink<C>>;
static_assert(lb::TraitsCodes<Traits<lb::D3Q19, lb::LBGK, lb::Normal, lb::BulkStreamer,
                                     BflByHand>>::wall == HLB_WALL_BFL, "link factory");

// The actor type itself instantiates (no device is touched at compile time).
using Actor = lb::LBM<Traits<lb::D3Q19, lb::DHumieresMRT, lb::Normal, lb::BulkStreamer,
                             lb::SimpleBounceBack>>;
static_assert(std::is_same<Actor::LatticeType, lb::D3Q19>::value, "actor lattice");
static_assert(!std::is_copy_constructible<Actor>::value, "the actor owns device state");

int main() { return 0; }
