#include "wpbr_ctx.hpp"
namespace wpbr {
void ctx_release_render(wpbr_ctx*) {}
}
