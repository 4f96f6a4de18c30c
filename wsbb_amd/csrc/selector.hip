// Learned machine-pair selector (placeholder until the selector lands).
#include "context.h"

namespace wsbb {
void destroy_selector(wsbb_ctx* ctx) { (void)ctx; }
}  // namespace wsbb
