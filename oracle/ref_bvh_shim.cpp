// ORACLE - test infrastructure only.  Exposes the reference's Morton encoder
// (src/avbd/bvh_broadphase.cpp:19-36, an internal-linkage helper) by compiling that
// translation unit from the reference tree inside this one.
#include "avbd/bvh_broadphase.cpp"

extern "C" uint64_t ref_morton3D(float x, float y, float z, float lo, float scale)
{
    return avbd::morton3D(x, y, z, lo, scale);
}
