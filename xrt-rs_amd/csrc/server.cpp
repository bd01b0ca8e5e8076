// placeholder, replaced by the serving loop
#include "runtime.h"
extern "C" {
int xrtk_server_create(xrtk_ctx*, const XrtkServerCfg*, xrtk_server** out) { if (out) *out = nullptr; return XRTK_EINVAL; }
int xrtk_server_submit(xrtk_server*, const XrtkJob*, xrtk_ticket*) { return XRTK_EINVAL; }
int xrtk_server_poll(xrtk_server*, xrtk_ticket) { return XRTK_EINVAL; }
int xrtk_server_wait(xrtk_server*, xrtk_ticket, double, double*) { return XRTK_EINVAL; }
int xrtk_server_stats(xrtk_server*, XrtkServerStats*) { return XRTK_EINVAL; }
void xrtk_server_destroy(xrtk_server*) {}
}
