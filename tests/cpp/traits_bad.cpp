// Combinations the GPU timestep does not provide must not compile
// (tests/test_cpp_layer.py compiles this once per BAD case).
#include "hemelb_amd.hpp"

using namespace hemelb;

#if BAD == 1   // an iolet streamer as the wall boundary
using T = Traits<lb::D3Q15, lb::LBGK, lb::Normal, lb::BulkStreamer, lb::NashZerothOrderPressure>;
#elif BAD == 2 // a wall streamer as an iolet boundary
using T = Traits<lb::D3Q15, lb::LBGK, lb::Normal, lb::BulkStreamer, lb::SimpleBounceBack,
                 lb::SimpleBounceBack>;
#elif BAD == 3 // a wall streamer for the bulk sites
using T = Traits<lb::D3Q19, lb::LBGK, lb::Normal, lb::BouzidiFirdaousLallemand>;
#elif BAD == 4 // a lattice the library does not provide
struct D3Q27 { static constexpr int NUMVECTORS = 27; };
using T = Traits<D3Q27, lb::LBGK>;
#endif

int main() { return lb::TraitsCodes<T>::lattice; }
