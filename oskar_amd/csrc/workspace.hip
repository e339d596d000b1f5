/* Per-(kind, device, stream) device workspaces (see workspace.h). */
#include <mutex>

#include "workspace.h"

namespace oskar_amd {

namespace {

struct Entry
{
    int kind, dev;
    hipStream_t stream;
    void* ptr;
    size_t size;
};

constexpr int MAX_ENTRIES = 1024;
Entry g_entries[MAX_ENTRIES];
int g_count = 0;
std::mutex g_lock;

} /* namespace */

void* ws_acquire(int kind, int dev, hipStream_t stream, size_t bytes,
        hipError_t* err)
{
    std::lock_guard<std::mutex> guard(g_lock);
    Entry* e = nullptr;
    for (int i = 0; i < g_count; ++i)
    {
        Entry& x = g_entries[i];
        if (x.kind == kind && x.dev == dev && x.stream == stream)
        {
            e = &x;
            break;
        }
    }
    if (!e)
    {
        if (g_count == MAX_ENTRIES)
        {
            *err = hipErrorOutOfMemory;
            return nullptr;
        }
        e = &g_entries[g_count++];
        e->kind = kind;
        e->dev = dev;
        e->stream = stream;
        e->ptr = nullptr;
        e->size = 0;
    }
    if (e->size < bytes)
    {
        if (e->ptr) (void) hipFree(e->ptr);
        e->ptr = nullptr;
        e->size = 0;
        *err = hipMalloc(&e->ptr, bytes > 0 ? bytes : 1);
        if (*err != hipSuccess)
        {
            e->ptr = nullptr;
            return nullptr;
        }
        e->size = bytes;
    }
    *err = hipSuccess;
    return e->ptr;
}

} /* namespace oskar_amd */
