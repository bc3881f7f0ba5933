// Host data buffer handed between the user and a backend (same public
// interface as OCTproEngine's include/iobuffer.h:10-34).  Buffers owned by
// the HIP backend hold pinned (page-locked) memory for asynchronous copies.
#ifndef OPE_IOBUFFER_H
#define OPE_IOBUFFER_H

#include <cstddef>

#include "export.h"
#include "types.h"

namespace ope {

namespace detail {
struct IOBufferAccess;
}

class OPE_API IOBuffer {
public:
    using DataType = ope::DataType;

    IOBuffer();
    ~IOBuffer();

    bool allocateMemory(size_t sizeInBytes);
    void releaseMemory();

    void setDataType(DataType type);
    DataType getDataType() const;
    int getBitDepth() const;

    void* getDataPointer();
    const void* getDataPointer() const;

    size_t getSizeInBytes() const;

private:
    // lets a backend adopt pinned memory; adds no member, so the layout is
    // the reference's three members
    friend struct detail::IOBufferAccess;
    void* dataPtr;
    size_t sizeInBytes;
    DataType dataType;
};

}  // namespace ope

#endif  // OPE_IOBUFFER_H
