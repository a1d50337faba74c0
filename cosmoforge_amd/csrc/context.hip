// Context lifetime, error reporting and the growable device workspace.
#include "cfmi_internal.h"

#include <cmath>
#include <vector>

int cfmi_fail(cfmi_ctx* c, int code, const char* fmt, ...)
{
    if (c) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

void* cfmi_workspace(cfmi_ctx* c, size_t bytes)
{
    if (bytes <= c->work_bytes) return c->d_work;
    // Stream-ordered: work already queued on the old buffer completes first.
    if (c->d_work) hipFreeAsync(c->d_work, c->stream);
    c->d_work = nullptr;
    c->work_bytes = 0;
    size_t want = bytes + bytes / 4;
    if (hipMallocAsync(&c->d_work, want, c->stream) != hipSuccess) {
        c->d_work = nullptr;
        return nullptr;
    }
    c->work_bytes = want;
    return c->d_work;
}

extern "C" {

int cfmi_version(void) { return 10000; }

int cfmi_create(int device, void* stream, cfmi_ctx** out)
{
    if (!out) return CFMI_EINVAL;
    *out = nullptr;
    cfmi_ctx* c = new cfmi_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return CFMI_EHIP;
    }
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return CFMI_EHIP;
        }
        c->own_stream = true;
    }
    if (hipMalloc((void**)&c->d_info, 64) != hipSuccess ||
        hipMalloc((void**)&c->d_red, CFMI_RED_SLOTS * sizeof(double)) != hipSuccess) {
        cfmi_destroy(c);
        return CFMI_ENOMEM;
    }
    if (hipMemset(c->d_info, 0, 64) != hipSuccess) {
        cfmi_destroy(c);
        return CFMI_EHIP;
    }
    *out = c;
    return CFMI_OK;
}

int cfmi_destroy(cfmi_ctx* c)
{
    if (!c) return CFMI_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& l : c->s8_launches) {
        hipEventDestroy(l.start);
        hipEventDestroy(l.stop);
    }
    for (auto e : c->s8_event_pool) hipEventDestroy(e);
    if (c->aux_stream) hipStreamSynchronize(c->aux_stream);
    if (c->aux_work) hipFree(c->aux_work);
    if (c->d_panel) hipFree(c->d_panel);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->aux_stream) hipStreamDestroy(c->aux_stream);
    if (c->d_work) hipFree(c->d_work);
    if (c->d_info) hipFree(c->d_info);
    if (c->d_red) hipFree(c->d_red);
    if (c->d_coef) hipFree(c->d_coef);
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    delete c;
    return CFMI_OK;
}

const char* cfmi_last_error(const cfmi_ctx* c) { return c ? c->err.c_str() : "null context"; }

int cfmi_synchronize(cfmi_ctx* c)
{
    if (!c) return CFMI_EINVAL;
    CFMI_HIP(c, hipStreamSynchronize(c->stream));
    return CFMI_OK;
}

}  // extern "C"
