// Internal: lets a backend give an IOBuffer memory it allocated itself
// (pinned host memory) together with the function that releases it.
#ifndef OPE_IOBUFFER_ACCESS_H
#define OPE_IOBUFFER_ACCESS_H

#include <cstddef>

#include "iobuffer.h"

namespace ope {
namespace detail {

struct IOBufferAccess {
    static void adopt(IOBuffer& b, void* ptr, size_t bytes, void (*release)(void*));
};

}  // namespace detail
}  // namespace ope

#endif  // OPE_IOBUFFER_ACCESS_H
