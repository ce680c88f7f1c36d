"""A restatement of the reference's SVG drawing (save_as_svg,
save_particle_info.cpp:130-270) from a downloaded state, with std::format's
"{}" float formatting (std::to_chars: the shortest digits that read back to
the same float, printed fixed or scientific, whichever is shorter, fixed on
a tie). Test infrastructure."""
import numpy as np

F = np.float32


def fmt32(v) -> str:
    v = F(v)
    if v == 0:
        return "-0" if np.signbit(v) else "0"
    sign = "-" if v < 0 else ""
    a = F(abs(v))
    sci = np.format_float_scientific(a, unique=True, trim="-")  # e.g. '1.2345679e+08', '1e-07'
    mant, exp = sci.split("e")
    e = int(exp)
    digits = mant.replace(".", "")
    # scientific: d[.ddd]e+XX with at least two exponent digits
    s_txt = digits[0] + ("." + digits[1:] if len(digits) > 1 else "") + "e" + ("-" if e < 0 else "+") + \
        f"{abs(e):02d}"
    # fixed: the shortest fraction that reads back; an integer value prints all its digits exactly
    if e >= len(digits) - 1:
        f_txt = str(int(a))
    elif e >= 0:
        f_txt = digits[:e + 1] + "." + digits[e + 1:]
    else:
        f_txt = "0." + "0" * (-e - 1) + digits
    return sign + (f_txt if len(f_txt) <= len(s_txt) else s_txt)


FILLS = ["0000ff", "ff0000", "00ff00", "ff8000", "ffff00", "ff00ff", "00ffff", "008000", "800000", "ffaec8"]


def svg_text(st: dict, boxes, view_min, view_max, render_scale, max_bd, color, render_boxes) -> str:
    n = len(st["radius"])
    include_boxes = len(boxes) > 0 and render_boxes
    include_kernels = color == 3
    gradient, count = (True, 0) if color == 1 else ((False, 10) if color == 2 else (False, 2))
    vmin, vmax = [F(x) for x in view_min], [F(x) for x in view_max]
    stroke = F(min(F(vmax[0] - vmin[0]), F(vmax[1] - vmin[1])) / F(1000.0))
    pc = np.zeros(n, np.int64)
    if color == 2:
        for i, (s, t) in enumerate(zip(st["t_source"], st["t_target"])):
            pc[s] = pc[t] = (i % (count - 1)) + 1
    elif color != 1:
        pc = (st["boundariness"] >= F(1.0)).astype(np.int64)
    orig = ""
    for c in range(count):
        orig += f'<circle cx="0" cy="0" r="{fmt32(render_scale)}" id="p{c}" style="fill:#{FILLS[c]};stroke-width:1" />'
    if include_kernels:
        orig += ('<circle cx="0" cy="0" r="1" id="k" style="fill:none;stroke:#ff7f00;stroke-width:'
                 f'{fmt32(stroke)};vector-effect:non-scaling-stroke" />')
    if include_boxes:
        orig += '<rect x="0" y="0" width="1" height="1" id="box" style="fill:#bffeff;stroke-width:1" />'
    c2 = F(F(F(max_bd) * F(262144.0)) * F(0.8))
    parts, kers = [], []
    for i in range(n):
        px = F(F(st["position"][i, 0]) / F(262144.0))
        py = -F(F(st["position"][i, 1]) / F(262144.0))
        rad, ker = st["radius"][i], st["kernel_width"][i]
        if gradient:
            a = F(F(F(st["boundary_distance"][i]) - F(0.0)) / F(c2 - F(0.0)))
            a = min(max(a, F(0.0)), F(1.0))
            col = []
            for x, y in zip((0.0, 0.0, 1.0), (0.62, 0.96, 0.83)):
                m = F(F(F(x) * F(F(1.0) - a)) + F(F(y) * a))
                col.append(int(F(m * F(255.0))))
            parts.append(f'<circle cx="{fmt32(px)}" cy="{fmt32(py)}" r="{fmt32(F(rad * F(render_scale)))}" '
                         f'style="fill:#{col[0]:02x}{col[1]:02x}{col[2]:02x};stroke-width:1" />')
        else:
            parts.append(f'<use transform="matrix({fmt32(rad)},0,0,{fmt32(rad)},{fmt32(px)},{fmt32(py)})" '
                         f'xlink:href="#p{pc[i]}" />')
        if include_kernels:
            kers.append(f'<use transform="matrix({fmt32(ker)},0,0,{fmt32(ker)},{fmt32(px)},{fmt32(py)})" '
                        'xlink:href="#k" />')
    svg = ""
    if orig:
        svg += f'<g id="originals" style="display:none">{orig}</g>'
    if include_boxes:
        b = ""
        for lo, hi in boxes:
            lo, hi = [F(x) for x in lo], [F(x) for x in hi]
            b += (f'<use transform="matrix({fmt32(hi[0] - lo[0])},0,0,{fmt32(hi[1] - lo[1])},{fmt32(lo[0])},'
                  f'{fmt32(-hi[1])})" xlink:href="#box" />')
        svg += f'<g id="boxes">{b}</g>'
    if include_kernels:
        svg += f'<g id="kernels">{"".join(kers)}</g>'
    svg += f'<g id="particles">{"".join(parts)}</g>'
    return (f'<svg viewBox="{fmt32(vmin[0])} {fmt32(-vmax[1])} {fmt32(F(vmax[0] - vmin[0]))} '
            f'{fmt32(F(vmax[1] - vmin[1]))}" version="1.1" xmlns="http://www.w3.org/2000/svg" '
            'xmlns:xlink="http://www.w3.org/1999/xlink" xmlns:svg="http://www.w3.org/2000/svg">'
            f'{svg}</svg>')
