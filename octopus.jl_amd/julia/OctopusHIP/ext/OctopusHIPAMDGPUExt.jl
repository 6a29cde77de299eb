#
# AMDGPU.jl glue of OctopusHIP: results as ROCArrays over the library's
# memory, and every call ordered on the task-local HIP stream.
#
module OctopusHIPAMDGPUExt

using AMDGPU
using OctopusHIP

# The raw hipStream_t of the caller's task-local stream: the library runs its
# kernels there, so they are ordered with the caller's own work (no host wait).
OctopusHIP._current_stream() = Ptr{Cvoid}(AMDGPU.stream().stream)

# A ROCArray over library-owned memory (no copy, not freed by Julia).  The
# finalizer closes over `owner` (the TNS), so the TNS - whose finalizer
# destroys the context and its buffers - stays reachable as long as the array
# does; a buffer that a later, larger result replaces is kept by the library
# until trim!/destruction, so the array stays valid memory.
function OctopusHIP._device_wrap(ptr::Ptr{T}, dims, owner) where {T}
    A = unsafe_wrap(ROCArray, ptr, dims; own=false)
    finalizer(_ -> (owner; nothing), A)
    return A
end

end # module
