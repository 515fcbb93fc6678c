// Octree state (filled in by the tree build).
#include "lpm_internal.hpp"

namespace lpm {
struct Tree {};
void tree_free(Tree* t) { delete t; }
}  // namespace lpm
