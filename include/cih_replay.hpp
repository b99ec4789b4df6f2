// cih_replay.hpp -- replay a command stream (cih_stream.h) into any object
// that exposes the canvas_ity::canvas public API (canvas_ity.hpp:177-1148).
//
// The same template drives the HIP canvas (canvas_ity_hip.hpp), the serving
// loop's request decoder and, in the test harness, the reference canvas, so a
// stream renders through each of them from exactly the same sequence of API
// calls.  Enum-typed arguments are converted with the canvas' own enum types.
#ifndef CIH_REPLAY_HPP
#define CIH_REPLAY_HPP

#include "cih_stream.h"

#include <cstddef>
#include <string>
#include <vector>

namespace cih {

struct stream_image {
    const unsigned char *texels;
    int width, height;
};

// Enum types are passed explicitly: BrushType (fill_style/stroke_style),
// RepetitionType, CompositeType, CapType, JoinType.
template <class Canvas, class BrushType, class RepetitionType,
          class CompositeType, class CapType, class JoinType>
struct replayer {
    Canvas &canvas;
    std::vector<stream_image> images;
    std::vector<float> dash;
    std::string string;

    explicit replayer(Canvas &target)
        : canvas(target), images(CIH_STREAM_MAX_IMAGES) {}

    static float f(const uint32_t *args, int index) { return cih_w2f(args[index]); }
    static int i(const uint32_t *args, int index) { return (int)(int32_t)args[index]; }

    // A text record's bytes as a C string (they end at a NUL, if any).
    const char *text(const uint32_t *args) {
        const char *bytes = reinterpret_cast<const char *>(args + 4);
        string.assign(bytes, bytes + i(args, 0));
        return string.c_str();
    }

    const stream_image *image(int id) const {
        if (id < 0 || id >= CIH_STREAM_MAX_IMAGES || !images[(size_t)id].texels)
            return 0;
        return &images[(size_t)id];
    }

    // Applies one record.  Returns 0, or -1 if the record names an undefined
    // image (the stream is then malformed).
    int apply(const cih_record &r) {
        const uint32_t *a = r.args;
        switch (r.op) {
        case CIH_OP_SCALE: canvas.scale(f(a, 0), f(a, 1)); break;
        case CIH_OP_ROTATE: canvas.rotate(f(a, 0)); break;
        case CIH_OP_TRANSLATE: canvas.translate(f(a, 0), f(a, 1)); break;
        case CIH_OP_TRANSFORM:
            canvas.transform(f(a, 0), f(a, 1), f(a, 2), f(a, 3), f(a, 4), f(a, 5));
            break;
        case CIH_OP_SET_TRANSFORM:
            canvas.set_transform(f(a, 0), f(a, 1), f(a, 2), f(a, 3), f(a, 4),
                                 f(a, 5));
            break;
        case CIH_OP_SET_GLOBAL_ALPHA: canvas.set_global_alpha(f(a, 0)); break;
        case CIH_OP_SET_COMPOSITE_OP:
            canvas.global_composite_operation = static_cast<CompositeType>(i(a, 0));
            break;
        case CIH_OP_SET_SHADOW_COLOR:
            canvas.set_shadow_color(f(a, 0), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_SET_SHADOW_BLUR: canvas.set_shadow_blur(f(a, 0)); break;
        case CIH_OP_SET_SHADOW_OFFSET:
            canvas.shadow_offset_x = f(a, 0);
            canvas.shadow_offset_y = f(a, 1);
            break;
        case CIH_OP_SET_LINE_WIDTH: canvas.set_line_width(f(a, 0)); break;
        case CIH_OP_SET_LINE_CAP: canvas.line_cap = static_cast<CapType>(i(a, 0)); break;
        case CIH_OP_SET_LINE_JOIN:
            canvas.line_join = static_cast<JoinType>(i(a, 0));
            break;
        case CIH_OP_SET_MITER_LIMIT: canvas.set_miter_limit(f(a, 0)); break;
        case CIH_OP_SET_LINE_DASH: {
            int count = i(a, 0);
            if (count < 0) {
                canvas.set_line_dash(0, 0);
                break;
            }
            dash.assign((size_t)count + 1, 0.0f);
            for (int k = 0; k < count; ++k)
                dash[(size_t)k] = f(a, 1 + k);
            canvas.set_line_dash(&dash[0], count);
            break;
        }
        case CIH_OP_SET_LINE_DASH_OFFSET: canvas.line_dash_offset = f(a, 0); break;
        case CIH_OP_SET_COLOR:
            canvas.set_color(static_cast<BrushType>(i(a, 0)), f(a, 1), f(a, 2),
                             f(a, 3), f(a, 4));
            break;
        case CIH_OP_SET_LINEAR_GRADIENT:
            canvas.set_linear_gradient(static_cast<BrushType>(i(a, 0)), f(a, 1),
                                       f(a, 2), f(a, 3), f(a, 4));
            break;
        case CIH_OP_SET_RADIAL_GRADIENT:
            canvas.set_radial_gradient(static_cast<BrushType>(i(a, 0)), f(a, 1),
                                       f(a, 2), f(a, 3), f(a, 4), f(a, 5), f(a, 6));
            break;
        case CIH_OP_ADD_COLOR_STOP:
            canvas.add_color_stop(static_cast<BrushType>(i(a, 0)), f(a, 1), f(a, 2),
                                  f(a, 3), f(a, 4), f(a, 5));
            break;
        case CIH_OP_SET_PATTERN: {
            const stream_image *img = image(i(a, 1));
            if (!img)
                return -1;
            canvas.set_pattern(static_cast<BrushType>(i(a, 0)), img->texels,
                               img->width, img->height, img->width * 4,
                               static_cast<RepetitionType>(i(a, 2)));
            break;
        }
        case CIH_OP_BEGIN_PATH: canvas.begin_path(); break;
        case CIH_OP_MOVE_TO: canvas.move_to(f(a, 0), f(a, 1)); break;
        case CIH_OP_CLOSE_PATH: canvas.close_path(); break;
        case CIH_OP_LINE_TO: canvas.line_to(f(a, 0), f(a, 1)); break;
        case CIH_OP_QUADRATIC_CURVE_TO:
            canvas.quadratic_curve_to(f(a, 0), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_BEZIER_CURVE_TO:
            canvas.bezier_curve_to(f(a, 0), f(a, 1), f(a, 2), f(a, 3), f(a, 4),
                                   f(a, 5));
            break;
        case CIH_OP_ARC_TO:
            canvas.arc_to(f(a, 0), f(a, 1), f(a, 2), f(a, 3), f(a, 4));
            break;
        case CIH_OP_ARC:
            canvas.arc(f(a, 0), f(a, 1), f(a, 2), f(a, 3), f(a, 4), i(a, 5) != 0);
            break;
        case CIH_OP_RECTANGLE:
            canvas.rectangle(f(a, 0), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_FILL: canvas.fill(); break;
        case CIH_OP_STROKE: canvas.stroke(); break;
        case CIH_OP_CLIP:
            if (clip(canvas, 0) != 0)
                return -1;
            break;
        case CIH_OP_CLEAR_RECTANGLE:
            canvas.clear_rectangle(f(a, 0), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_FILL_RECTANGLE:
            canvas.fill_rectangle(f(a, 0), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_STROKE_RECTANGLE:
            canvas.stroke_rectangle(f(a, 0), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_DRAW_IMAGE: {
            const stream_image *img = image(i(a, 0));
            if (!img)
                return -1;
            canvas.draw_image(img->texels, img->width, img->height, img->width * 4,
                              f(a, 1), f(a, 2), f(a, 3), f(a, 4));
            break;
        }
        case CIH_OP_SAVE: canvas.save(); break;
        case CIH_OP_RESTORE: canvas.restore(); break;
        case CIH_OP_DEFINE_IMAGE: {
            stream_image &img = images[(size_t)i(a, 0)];
            img.texels = reinterpret_cast<const unsigned char *>(a + 3);
            img.width = i(a, 1);
            img.height = i(a, 2);
            break;
        }
        case CIH_OP_PUT_IMAGE_DATA: {
            const stream_image *img = image(i(a, 0));
            if (!img)
                return -1;
            canvas.put_image_data(img->texels, img->width, img->height,
                                  img->width * 4, i(a, 1), i(a, 2));
            break;
        }
        case CIH_OP_SET_FILL_RULE: set_fill_rule(canvas, i(a, 0), 0); break;
        case CIH_OP_SET_FONT: {
            int bytes = i(a, 0);
            const unsigned char *font = reinterpret_cast<const unsigned char *>(a + 2);
            canvas.set_font(bytes ? font : 0, bytes, f(a, 1));
            break;
        }
        case CIH_OP_FILL_TEXT:
            canvas.fill_text(text(a), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_STROKE_TEXT:
            canvas.stroke_text(text(a), f(a, 1), f(a, 2), f(a, 3));
            break;
        case CIH_OP_SET_TEXT_ALIGN:
            canvas.text_align = static_cast<decltype(canvas.text_align)>(i(a, 0));
            break;
        case CIH_OP_SET_TEXT_BASELINE:
            canvas.text_baseline = static_cast<decltype(canvas.text_baseline)>(i(a, 0));
            break;
        default: return -1;
        }
        return 0;
    }

    // Canvases without clip() reject the record.
    template <class C>
    static auto clip(C &c, int) -> decltype(c.clip(), int()) {
        c.clip();
        return 0;
    }
    template <class C>
    static int clip(C &, long) { return -1; }

    // The fill rule is an extension: canvases without it ignore the record.
    template <class C>
    static auto set_fill_rule(C &c, int rule, int) -> decltype(c.set_fill_rule(rule), void()) {
        c.set_fill_rule(rule);
    }
    template <class C>
    static void set_fill_rule(C &, int, long) {}

    // Replays every record.  Returns 0, or -1 for a malformed stream (records
    // before the malformed one have been applied).
    int run(const void *stream, size_t bytes) {
        cih_reader reader;
        if (cih_reader_init(&reader, stream, bytes) != 0)
            return -1;
        cih_record record;
        int status;
        while ((status = cih_reader_next(&reader, &record)) == 1)
            if (apply(record) != 0)
                return -1;
        return status;
    }
};

} // namespace cih

#endif // CIH_REPLAY_HPP
