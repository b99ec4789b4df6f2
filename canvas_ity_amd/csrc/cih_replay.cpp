// cih_replay.cpp -- cih_canvas_replay: a command stream (cih_stream.h) played
// through canvas_ity_hip.hpp's canvas API onto an existing device canvas.
// This is how batch and serving requests, and the test harness, drive the
// same host state machine that a C++ caller of the header uses.
#include "canvas_ity_hip.hpp"
#include "cih_replay.hpp"

int cih_internal_fail(int status, const char *message);

extern "C" CIH_API int cih_canvas_replay(cih_canvas *target, const void *stream, size_t bytes) {
    if (!target)
        return cih_internal_fail(CIH_ERROR_INVALID, "canvas is NULL");
    cih_reader reader;
    if (cih_reader_init(&reader, stream, bytes) != 0)
        return cih_internal_fail(CIH_ERROR_MALFORMED, "malformed command stream header");
    int width = 0, height = 0;
    cih_canvas_size(target, &width, &height);
    if (reader.width != width || reader.height != height)
        return cih_internal_fail(CIH_ERROR_INVALID, "stream size differs from the canvas");
    int status;
    {
        canvas_ity_hip::canvas canvas(target);
        cih::replayer<canvas_ity_hip::canvas, canvas_ity_hip::brush_type,
                      canvas_ity_hip::repetition_style, canvas_ity_hip::composite_operation,
                      canvas_ity_hip::cap_style, canvas_ity_hip::join_style>
            replay(canvas);
        status = replay.run(stream, bytes);
    } // the canvas submits its pending packets as it goes out of scope
    if (status != 0)
        return cih_internal_fail(CIH_ERROR_MALFORMED, "malformed command stream record");
    return CIH_OK;
}
