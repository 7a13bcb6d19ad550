// tbk_prepared.h -- plan creation in two parts, for the request server: the
// host set-up of a structure (neighbour list, bond integrals, SK blocks) on
// any host thread, then the device part (upload, workspace) on a stream the
// caller lends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "tbk.h"

struct tbk_prepared;

int tbk_prepare(const tbk_structure_desc *st, tbk_prepared **out);
void tbk_prepared_free(tbk_prepared *p);
// states per k of the prepared structure
int tbk_prepared_nv(const tbk_prepared *p);
// device bytes its set-up arrays take in a plan
size_t tbk_prepared_bytes(const tbk_prepared *p);
// A group plan (all k of all structures in one batch) on `stream`, which the
// plan uses and does not destroy; the prepared structures stay the caller's.
// The matrix order is at least `order` (within the solver's limit), so that
// a structure's results do not depend on the others in the group.
int tbk_plan_create_prepared(int device, int nstruct, const tbk_prepared *const *ps, const int *nk,
                             const double *kpts, const double *weights, int precision, hipStream_t stream,
                             int order, tbk_plan **out);
