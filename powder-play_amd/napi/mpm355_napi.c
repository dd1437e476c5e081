/*
 * Node-API binding of libmpm355 (include/mpm355.h): the native runtime a
 * TypeScript MpmMlsGpuSolver implementation calls instead of WebGPU
 * (src/lib/mpm/kernels/solver.ts:106-144).  Every function forwards to one
 * C-ABI entry point; argument errors become RangeError/TypeError before any
 * device work (solver.ts:1414-1441) and device failures become Error with a
 * `code` property.  readCompact returns a Promise resolved from a worker
 * thread, like the reference's mapAsync readback (solver.ts:2388-2512).
 *
 * The C ABI asks callers to serialise the calls on one context.  Each context
 * object therefore carries a mutex: every call holds it for the duration of
 * its C-ABI call, and so does the readCompact worker while it waits for the
 * GPU.  JS may keep submitting work while a readback is pending (the
 * reference allows that); such a call waits for the readback's wait to end.
 */
#include <node_api.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpm355.h"

#define CHECK(call)                                                          \
    do {                                                                     \
        if ((call) != napi_ok) {                                             \
            napi_throw_error(env, NULL, "Node-API call failed: " #call);     \
            return NULL;                                                     \
        }                                                                    \
    } while (0)

static napi_value throw_code(napi_env env, int code, mpm_ctx* ctx) {
    const char* msg = mpm_last_error(ctx);
    if (code == MPM_EINVAL) {
        napi_throw_range_error(env, "MPM_EINVAL", msg);
    } else {
        const char* name = code == MPM_EDEVICE ? "MPM_EDEVICE" : code == MPM_ESTATE ? "MPM_ESTATE" : "MPM_ECAPACITY";
        napi_throw_error(env, name, msg);
    }
    return NULL;
}

/* A solver context and the lock that serialises the calls on it. */
typedef struct {
    mpm_ctx* ctx;
    pthread_mutex_t lock;
} ctx_box;

static void finalize_ctx(napi_env env, void* data, void* hint) {
    (void)env;
    (void)hint;
    ctx_box* b = (ctx_box*)data;
    mpm_destroy(b->ctx);
    pthread_mutex_destroy(&b->lock);
    free(b);
}

/* The context of argument 0, or NULL after a thrown TypeError. */
static ctx_box* ctx_arg(napi_env env, napi_value v) {
    napi_valuetype t;
    void* p = NULL;
    if (napi_typeof(env, v, &t) != napi_ok || t != napi_external || napi_get_value_external(env, v, &p) != napi_ok ||
        !p) {
        napi_throw_type_error(env, NULL, "expected a solver context");
        return NULL;
    }
    return (ctx_box*)p;
}

/* Throw for a call that failed under the context's lock, then release it
 * (the message is read while the lock is still held). */
static napi_value fail_unlock(napi_env env, int code, ctx_box* b) {
    throw_code(env, code, b->ctx);
    pthread_mutex_unlock(&b->lock);
    return NULL;
}

/* Run one C-ABI call under the context's lock; throw on failure. */
#define LOCKED_CALL(box, call)                                     \
    do {                                                           \
        pthread_mutex_lock(&(box)->lock);                          \
        const int code_ = (call);                                  \
        if (code_ != MPM_OK) return fail_unlock(env, code_, box);  \
        pthread_mutex_unlock(&(box)->lock);                        \
    } while (0)

static int get_number(napi_env env, napi_value obj, const char* key, double* out, int required) {
    bool has = false;
    napi_value v;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok) return 0;
    if (!has) {
        if (required) {
            napi_throw_type_error(env, NULL, key);
            return 0;
        }
        return 1;
    }
    if (napi_get_named_property(env, obj, key, &v) != napi_ok) return 0;
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_boolean) {
        bool b;
        napi_get_value_bool(env, v, &b);
        *out = b ? 1.0 : 0.0;
        return 1;
    }
    if (t != napi_number || napi_get_value_double(env, v, out) != napi_ok) {
        napi_throw_type_error(env, NULL, key);
        return 0;
    }
    return 1;
}

/* A typed array argument as (data, element count), or 0 after a TypeError. */
static int typed_arg(napi_env env, napi_value v, napi_typedarray_type want, void** data, size_t* count,
                     int allow_null) {
    napi_valuetype vt;
    napi_typeof(env, v, &vt);
    if (allow_null && (vt == napi_null || vt == napi_undefined)) {
        *data = NULL;
        *count = 0;
        return 1;
    }
    bool is = false;
    napi_is_typedarray(env, v, &is);
    napi_typedarray_type t;
    napi_value ab;
    size_t off;
    if (!is || napi_get_typedarray_info(env, v, &t, count, data, &ab, &off) != napi_ok || t != want) {
        napi_throw_type_error(env, NULL, "expected a typed array of the documented element type");
        return 0;
    }
    return 1;
}

/* create(config, packed: Uint32Array, objectIds: Uint32Array | null, atomic: boolean) */
static napi_value create(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 4) {
        napi_throw_type_error(env, NULL, "create(config, packed, objectIds, atomic)");
        return NULL;
    }
    double cx = 0, cy = 0, cs = 0, pm = 0, pv = 0, wf = 0, capacity = 0, device = -1, slots = 0, keep = 0;
    if (!get_number(env, argv[0], "cellsX", &cx, 1) || !get_number(env, argv[0], "cellsY", &cy, 1) ||
        !get_number(env, argv[0], "cellSize", &cs, 1) || !get_number(env, argv[0], "particleMass", &pm, 1) ||
        !get_number(env, argv[0], "particleVolume", &pv, 1) || !get_number(env, argv[0], "wallFriction", &wf, 0) ||
        !get_number(env, argv[0], "capacity", &capacity, 0) || !get_number(env, argv[0], "device", &device, 0) ||
        !get_number(env, argv[0], "bucketSlots", &slots, 0) || !get_number(env, argv[0], "keepFixedGrid", &keep, 0))
        return NULL;
    void *packed, *ids;
    size_t words, nid;
    if (!typed_arg(env, argv[1], napi_uint32_array, &packed, &words, 0)) return NULL;
    if (!typed_arg(env, argv[2], napi_uint32_array, &ids, &nid, 1)) return NULL;
    bool atomic = false;
    CHECK(napi_get_value_bool(env, argv[3], &atomic));
    if (words % MPM_PARTICLE_WORDS != 0) {
        napi_throw_range_error(env, "MPM_EINVAL", "packed particles must be a multiple of 20 words");
        return NULL;
    }
    const uint32_t n = (uint32_t)(words / MPM_PARTICLE_WORDS);
    if (ids && nid != n) {
        napi_throw_range_error(env, "MPM_EINVAL", "objectIds must hold one word per particle");
        return NULL;
    }
    mpm_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.cells_x = (uint32_t)cx;
    cfg.cells_y = (uint32_t)cy;
    cfg.cell_size = (float)cs;
    cfg.particle_mass = (float)pm;
    cfg.particle_volume = (float)pv;
    cfg.wall_friction = (float)wf;
    cfg.capacity = capacity > 0 ? (uint32_t)capacity : n;
    cfg.p2g_mode = atomic ? MPM_P2G_FIXED_POINT : MPM_P2G_GATHER;
    cfg.flags = keep != 0 ? MPM_FLAG_KEEP_FIXED_GRID : 0u;
    cfg.device = (int32_t)device;
    cfg.bucket_slots = (uint32_t)slots;
    mpm_ctx* ctx = NULL;
    const int code = mpm_create(&cfg, (const uint32_t*)packed, (const uint32_t*)ids, n, &ctx);
    if (code != MPM_OK) return throw_code(env, code, NULL);
    ctx_box* box = (ctx_box*)calloc(1, sizeof *box);
    if (!box) {
        mpm_destroy(ctx);
        napi_throw_error(env, NULL, "out of memory");
        return NULL;
    }
    box->ctx = ctx;
    pthread_mutex_init(&box->lock, NULL);
    napi_value ext;
    if (napi_create_external(env, box, finalize_ctx, NULL, &ext) != napi_ok) {
        finalize_ctx(env, box, NULL);
        napi_throw_error(env, NULL, "Node-API call failed: napi_create_external");
        return NULL;
    }
    return ext;
}

/* stepBatch(ctx, params: Uint8Array(144) | ArrayBuffer(144), stepCount, measureSpeed) */
static napi_value step_batch(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 4 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    void* params = NULL;
    size_t len = 0;
    bool is_ab = false;
    napi_is_arraybuffer(env, argv[1], &is_ab);
    if (is_ab) {
        CHECK(napi_get_arraybuffer_info(env, argv[1], &params, &len));
    } else if (!typed_arg(env, argv[1], napi_uint8_array, &params, &len, 0)) {
        return NULL;
    }
    if (len != MPM_PARAMS_BYTES) {
        napi_throw_range_error(env, "MPM_EINVAL", "step parameters must be 144 bytes");
        return NULL;
    }
    uint32_t steps = 0;
    bool measure = false;
    CHECK(napi_get_value_uint32(env, argv[2], &steps));
    CHECK(napi_get_value_bool(env, argv[3], &measure));
    LOCKED_CALL(box, mpm_step_batch(box->ctx, params, steps, measure ? 1u : 0u));
    return NULL;
}

static napi_value compact_object(napi_env env, float speed, uint32_t completed, const uint32_t status[8]) {
    napi_value out, v, ab, arr;
    void* data;
    if (napi_create_object(env, &out) != napi_ok) return NULL;
    napi_create_double(env, speed, &v);
    napi_set_named_property(env, out, "maxSpeed", v);
    napi_create_uint32(env, completed, &v);
    napi_set_named_property(env, out, "completed", v);
    napi_create_arraybuffer(env, 32, &data, &ab);
    memcpy(data, status, 32);
    napi_create_typedarray(env, napi_uint32_array, 8, ab, 0, &arr);
    napi_set_named_property(env, out, "status", arr);
    return out;
}

/* readCompactSync(ctx) -> { maxSpeed, completed, status: Uint32Array(8) } */
static napi_value read_compact_sync(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    float speed = 0;
    uint32_t completed = 0, status[8];
    LOCKED_CALL(box, mpm_read_compact(box->ctx, &speed, &completed, status));
    return compact_object(env, speed, completed, status);
}

typedef struct {
    napi_async_work work;
    napi_deferred deferred;
    napi_ref ctx_ref; /* keeps the context alive while the worker waits */
    ctx_box* box;
    float speed;
    uint32_t completed, status[8];
    int code;
    char message[512]; /* the library's message of a failed read */
} compact_job;

static void compact_exec(napi_env env, void* p) {
    (void)env;
    compact_job* j = (compact_job*)p;
    pthread_mutex_lock(&j->box->lock);
    j->code = mpm_read_compact(j->box->ctx, &j->speed, &j->completed, j->status);
    if (j->code != MPM_OK) snprintf(j->message, sizeof j->message, "%s", mpm_last_error(j->box->ctx));
    pthread_mutex_unlock(&j->box->lock);
}

static void compact_done(napi_env env, napi_status s, void* p) {
    compact_job* j = (compact_job*)p;
    (void)s;
    if (j->code != MPM_OK) {
        napi_value msg, err, code;
        napi_create_string_utf8(env, j->message, NAPI_AUTO_LENGTH, &msg);
        napi_create_string_utf8(env, j->code == MPM_EINVAL ? "MPM_EINVAL" : "MPM_EDEVICE", NAPI_AUTO_LENGTH, &code);
        napi_create_error(env, code, msg, &err);
        napi_reject_deferred(env, j->deferred, err);
    } else {
        napi_resolve_deferred(env, j->deferred, compact_object(env, j->speed, j->completed, j->status));
    }
    napi_delete_reference(env, j->ctx_ref);
    napi_delete_async_work(env, j->work);
    free(j);
}

/* readCompact(ctx) -> Promise<{ maxSpeed, completed, status }> */
static napi_value read_compact(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], promise, name;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    compact_job* j = (compact_job*)calloc(1, sizeof *j);
    if (!j) {
        napi_throw_error(env, NULL, "out of memory");
        return NULL;
    }
    j->box = box;
    CHECK(napi_create_reference(env, argv[0], 1, &j->ctx_ref));
    CHECK(napi_create_promise(env, &j->deferred, &promise));
    CHECK(napi_create_string_utf8(env, "mpm_read_compact", NAPI_AUTO_LENGTH, &name));
    CHECK(napi_create_async_work(env, NULL, name, compact_exec, compact_done, j, &j->work));
    CHECK(napi_queue_async_work(env, j->work));
    return promise;
}

/* clearStatus(ctx) */
static napi_value clear_status(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    LOCKED_CALL(box, mpm_clear_status(box->ctx));
    return NULL;
}

/* counts(ctx) -> { particleCount, particleCapacity, nodeCount } */
static napi_value counts(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], out, v;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    uint32_t n = 0, cap = 0, nodes = 0;
    LOCKED_CALL(box, mpm_get_counts(box->ctx, &n, &cap, &nodes));
    CHECK(napi_create_object(env, &out));
    napi_create_uint32(env, n, &v);
    napi_set_named_property(env, out, "particleCount", v);
    napi_create_uint32(env, cap, &v);
    napi_set_named_property(env, out, "particleCapacity", v);
    napi_create_uint32(env, nodes, &v);
    napi_set_named_property(env, out, "nodeCount", v);
    return out;
}

/* readParticles(ctx, indices: Uint32Array | null) -> { records: Uint32Array, objectIds: Uint32Array } */
static napi_value read_particles(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2], out, rec_ab, ids_ab, rec, ids;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    void* idx = NULL;
    size_t nidx = 0;
    if (argc >= 2 && !typed_arg(env, argv[1], napi_uint32_array, &idx, &nidx, 1)) return NULL;
    uint32_t n = (uint32_t)nidx;
    pthread_mutex_lock(&box->lock);
    if (!idx) {
        uint32_t cap = 0, nodes = 0;
        const int code = mpm_get_counts(box->ctx, &n, &cap, &nodes);
        if (code != MPM_OK) return fail_unlock(env, code, box);
    }
    void *rd, *id;
    if (napi_create_arraybuffer(env, (size_t)n * MPM_PARTICLE_WORDS * 4, &rd, &rec_ab) != napi_ok ||
        napi_create_arraybuffer(env, (size_t)n * 4, &id, &ids_ab) != napi_ok) {
        pthread_mutex_unlock(&box->lock);
        napi_throw_error(env, NULL, "Node-API call failed: napi_create_arraybuffer");
        return NULL;
    }
    const int code = mpm_read_particles(box->ctx, (const uint32_t*)idx, n, (uint32_t*)rd, (uint32_t*)id);
    if (code != MPM_OK) return fail_unlock(env, code, box);
    pthread_mutex_unlock(&box->lock);
    CHECK(napi_create_typedarray(env, napi_uint32_array, (size_t)n * MPM_PARTICLE_WORDS, rec_ab, 0, &rec));
    CHECK(napi_create_typedarray(env, napi_uint32_array, n, ids_ab, 0, &ids));
    CHECK(napi_create_object(env, &out));
    napi_set_named_property(env, out, "records", rec);
    napi_set_named_property(env, out, "objectIds", ids);
    return out;
}

/* appendParticles(ctx, packed: Uint32Array, objectIds: Uint32Array | null) */
static napi_value append_particles(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 3 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    void *packed, *ids;
    size_t words, nid;
    if (!typed_arg(env, argv[1], napi_uint32_array, &packed, &words, 0)) return NULL;
    if (!typed_arg(env, argv[2], napi_uint32_array, &ids, &nid, 1)) return NULL;
    if (words % MPM_PARTICLE_WORDS != 0 || (ids && nid != words / MPM_PARTICLE_WORDS)) {
        napi_throw_range_error(env, "MPM_EINVAL", "packed particles and objectIds disagree");
        return NULL;
    }
    LOCKED_CALL(box, mpm_append_particles(box->ctx, (const uint32_t*)packed, (const uint32_t*)ids,
                                          (uint32_t)(words / MPM_PARTICLE_WORDS)));
    return NULL;
}

/* setInteraction(ctx, Float32Array(8)) */
static napi_value set_interaction(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 2 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    void* p;
    size_t n;
    if (!typed_arg(env, argv[1], napi_float32_array, &p, &n, 0)) return NULL;
    if (n != 8) {
        napi_throw_range_error(env, "MPM_EINVAL", "interaction uniform must be 8 floats");
        return NULL;
    }
    LOCKED_CALL(box, mpm_set_interaction(box->ctx, (const float*)p));
    return NULL;
}

/* removeParticlesInOutlet(ctx, Float32Array(4) {minX, minY, maxX, maxY}) -> removed count */
static napi_value remove_in_outlet(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2], out;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 2 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    void* p;
    size_t n;
    if (!typed_arg(env, argv[1], napi_float32_array, &p, &n, 0)) return NULL;
    if (n != 4) {
        napi_throw_range_error(env, "MPM_EINVAL", "outlet bounds must be 4 floats");
        return NULL;
    }
    uint32_t removed = 0;
    LOCKED_CALL(box, mpm_remove_in_outlet(box->ctx, (const float*)p, &removed));
    CHECK(napi_create_uint32(env, removed, &out));
    return out;
}

/* gatherOccupancy(ctx) -> Float32Array(2 * nodeCount): (water, oil) per node */
static napi_value gather_occupancy(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], ab, arr;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    uint32_t n = 0, cap = 0, nodes = 0;
    LOCKED_CALL(box, mpm_get_counts(box->ctx, &n, &cap, &nodes));
    void* data;
    CHECK(napi_create_arraybuffer(env, (size_t)nodes * 8, &data, &ab));
    LOCKED_CALL(box, mpm_gather_occupancy(box->ctx, (float*)data, (size_t)nodes * 2));
    CHECK(napi_create_typedarray(env, napi_float32_array, (size_t)nodes * 2, ab, 0, &arr));
    return arr;
}

/* pick(ctx, x, y, radius) -> Uint32Array(4) {index | 0xFFFFFFFF, x bits, y bits, objectId} */
static napi_value pick(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4], ab, arr;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 4 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    double v[3];
    for (int k = 0; k < 3; ++k)
        if (napi_get_value_double(env, argv[k + 1], &v[k]) != napi_ok) {
            napi_throw_type_error(env, NULL, "pick point and radius must be numbers");
            return NULL;
        }
    void* data;
    CHECK(napi_create_arraybuffer(env, 16, &data, &ab));
    LOCKED_CALL(box, mpm_pick(box->ctx, (float)v[0], (float)v[1], (float)v[2], (uint32_t*)data));
    CHECK(napi_create_typedarray(env, napi_uint32_array, 4, ab, 0, &arr));
    return arr;
}

/* clone(ctx, start, count, offsetX, offsetY, objectId) -> boolean (false when
 * the reference refuses: capacity, outlet removal, fracture bonds) */
static napi_value clone_object(napi_env env, napi_callback_info info) {
    size_t argc = 6;
    napi_value argv[6], out;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 6 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    uint32_t start = 0, count = 0, oid = 0;
    double ox = 0, oy = 0;
    if (napi_get_value_uint32(env, argv[1], &start) != napi_ok || napi_get_value_uint32(env, argv[2], &count) != napi_ok ||
        napi_get_value_double(env, argv[3], &ox) != napi_ok || napi_get_value_double(env, argv[4], &oy) != napi_ok ||
        napi_get_value_uint32(env, argv[5], &oid) != napi_ok) {
        napi_throw_type_error(env, NULL, "clone(ctx, start, count, offsetX, offsetY, objectId)");
        return NULL;
    }
    pthread_mutex_lock(&box->lock);
    const int code = mpm_clone(box->ctx, start, count, (float)ox, (float)oy, oid);
    if (code != MPM_OK && code != MPM_ECAPACITY && code != MPM_ESTATE) return fail_unlock(env, code, box);
    pthread_mutex_unlock(&box->lock);
    CHECK(napi_get_boolean(env, code == MPM_OK, &out));
    return out;
}

/* setDiagnostics(ctx, enable) / readDiagnostics(ctx) -> Float64Array(9) */
static napi_value set_diagnostics(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 2 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    bool on = false;
    CHECK(napi_get_value_bool(env, argv[1], &on));
    LOCKED_CALL(box, mpm_set_diagnostics(box->ctx, on ? 1u : 0u));
    return NULL;
}

static napi_value read_diagnostics(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], ab, arr;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    void* data;
    CHECK(napi_create_arraybuffer(env, 9 * sizeof(double), &data, &ab));
    LOCKED_CALL(box, mpm_read_diagnostics(box->ctx, (double*)data));
    CHECK(napi_create_typedarray(env, napi_float64_array, 9, ab, 0, &arr));
    return arr;
}

/* setFractureBonds(ctx, Uint32Array(6 * n)): the 24-byte bond records of
 * config.fractureBonds (solver.ts:929-935) */
static napi_value set_fracture_bonds(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 2 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    void* p;
    size_t words;
    if (!typed_arg(env, argv[1], napi_uint32_array, &p, &words, 0)) return NULL;
    if (words % 6 != 0) {
        napi_throw_range_error(env, "MPM_EINVAL", "bond records are 6 words each");
        return NULL;
    }
    LOCKED_CALL(box, mpm_set_fracture_bonds(box->ctx, p, (uint32_t)(words / 6)));
    return NULL;
}

/* readFractureBonds(ctx, count) -> Uint32Array(6 * count) */
static napi_value read_fracture_bonds(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2], ab, arr;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 2 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    uint32_t count = 0;
    if (napi_get_value_uint32(env, argv[1], &count) != napi_ok) {
        napi_throw_type_error(env, NULL, "count must be a number");
        return NULL;
    }
    void* data;
    CHECK(napi_create_arraybuffer(env, (size_t)count * 24, &data, &ab));
    LOCKED_CALL(box, mpm_read_fracture_bonds(box->ctx, data, count));
    CHECK(napi_create_typedarray(env, napi_uint32_array, (size_t)count * 6, ab, 0, &arr));
    return arr;
}

/* setTuning(ctx, {p2gTile, g2pTile, g2pThreads, bucketCapacity, tileQueue}) and
 * getTuning(ctx) -> the same object (mpm_tuning; no reference counterpart).
 * bucketCapacity and tileQueue may be left out (0: keep the capacity, the
 * automatic tile queue). */
static napi_value set_tuning(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 2 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    double v[5] = {0, 0, 0, 0, 0};
    const char* keys[5] = {"p2gTile", "g2pTile", "g2pThreads", "bucketCapacity", "tileQueue"};
    for (int k = 0; k < 5; ++k)
        if (!get_number(env, argv[1], keys[k], &v[k], k < 3)) return NULL;
    for (int k = 0; k < 5; ++k)
        if (!(v[k] >= 0 && v[k] <= 4294967295.0) || v[k] != (double)(uint32_t)v[k]) {
            napi_throw_range_error(env, "MPM_EINVAL", keys[k]);
            return NULL;
        }
    mpm_tuning t = {(uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], (uint32_t)v[3], (uint32_t)v[4]};
    LOCKED_CALL(box, mpm_set_tuning(box->ctx, &t));
    return NULL;
}

static napi_value get_tuning(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], out, x;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    ctx_box* box = argc >= 1 ? ctx_arg(env, argv[0]) : NULL;
    if (!box) return NULL;
    mpm_tuning t;
    LOCKED_CALL(box, mpm_get_tuning(box->ctx, &t));
    CHECK(napi_create_object(env, &out));
    const uint32_t vals[5] = {t.p2g_tile, t.g2p_tile, t.g2p_threads, t.bucket_capacity, t.tile_queue};
    const char* keys[5] = {"p2gTile", "g2pTile", "g2pThreads", "bucketCapacity", "tileQueue"};
    for (int k = 0; k < 5; ++k) {
        CHECK(napi_create_uint32(env, vals[k], &x));
        napi_set_named_property(env, out, keys[k], x);
    }
    return out;
}

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor p[] = {
        {"create", 0, create, 0, 0, 0, napi_enumerable, 0},
        {"stepBatch", 0, step_batch, 0, 0, 0, napi_enumerable, 0},
        {"readCompact", 0, read_compact, 0, 0, 0, napi_enumerable, 0},
        {"readCompactSync", 0, read_compact_sync, 0, 0, 0, napi_enumerable, 0},
        {"clearStatus", 0, clear_status, 0, 0, 0, napi_enumerable, 0},
        {"counts", 0, counts, 0, 0, 0, napi_enumerable, 0},
        {"readParticles", 0, read_particles, 0, 0, 0, napi_enumerable, 0},
        {"appendParticles", 0, append_particles, 0, 0, 0, napi_enumerable, 0},
        {"setInteraction", 0, set_interaction, 0, 0, 0, napi_enumerable, 0},
        {"removeParticlesInOutlet", 0, remove_in_outlet, 0, 0, 0, napi_enumerable, 0},
        {"pick", 0, pick, 0, 0, 0, napi_enumerable, 0},
        {"gatherOccupancy", 0, gather_occupancy, 0, 0, 0, napi_enumerable, 0},
        {"clone", 0, clone_object, 0, 0, 0, napi_enumerable, 0},
        {"setDiagnostics", 0, set_diagnostics, 0, 0, 0, napi_enumerable, 0},
        {"readDiagnostics", 0, read_diagnostics, 0, 0, 0, napi_enumerable, 0},
        {"setFractureBonds", 0, set_fracture_bonds, 0, 0, 0, napi_enumerable, 0},
        {"readFractureBonds", 0, read_fracture_bonds, 0, 0, 0, napi_enumerable, 0},
        {"setTuning", 0, set_tuning, 0, 0, 0, napi_enumerable, 0},
        {"getTuning", 0, get_tuning, 0, 0, 0, napi_enumerable, 0},
    };
    if (napi_define_properties(env, exports, sizeof p / sizeof p[0], p) != napi_ok) return NULL;
    return exports;
}

NAPI_MODULE(mpm355, init)
