// Learned preconditioner (placeholder until the message-passing network lands).
#include "pdg_ctx.hpp"

namespace pdg {

struct GnnState {};

void Ctx::free_gnn() {
  delete gnn;
  gnn = nullptr;
}

int gnn_apply(Ctx* c, const double* r, double* z) {
  (void)r; (void)z;
  return c->fail(PDG_E_UNSUPPORTED, "learned preconditioner not trained");
}

}  // namespace pdg
