/* ref_proto.c: TEST INFRASTRUCTURE. Exports the reference's own inline message check (src/common/vt_proto.h:181-191)
 * so the tests can pin vtg_msg_check against it. Built only by `make ref`, from the reference tree's header, into
 * oracle/_ref/. */
#include "vt_proto.h"

uint64_t ref_vt_msg_check(const struct vt_msg *m) { return vt_msg_check(m); }
unsigned ref_vt_msg_size(void) { return (unsigned)sizeof(struct vt_msg); }
