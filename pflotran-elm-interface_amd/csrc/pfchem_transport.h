// pfchem_transport.h -- internal interface between the context
// (pfchem_api.hip) and the OSRT transport step (pfchem_transport.hip).
#pragma once

#include "pfchem.h"

namespace pfchem {

struct TransportPlan;  // connection topology on the device (pfchem_transport.hip)
void transport_plan_free(TransportPlan* p);

// context accessors (pfchem_api.hip)
int ctx_device(const pfchem_ctx* ctx);
int ctx_naq(const pfchem_ctx* ctx);
TransportPlan** ctx_transport_plan(pfchem_ctx* ctx);
int ctx_fail(pfchem_ctx* ctx, int code, const char* msg);

}  // namespace pfchem
