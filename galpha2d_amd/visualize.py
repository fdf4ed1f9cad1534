"""Static picture of a result: the triangulation and the alpha complex at one rank, as SVG.

The reference's viewer (AlphaShape/Visualizer.cu, drawBasics.cu) is an interactive OpenGL window
that draws the DT and the alpha shape at the current alpha index (right mouse changes it).  This
writes the same picture for one rank to a file, from the classes that Context.classify (or
galphaClassify) returns: interior triangles filled, regular edges (the shape's boundary) bold,
singular edges thin, singular vertices as dots, the rest of the triangulation faint.
"""
import numpy as np

EXTERIOR, SINGULAR, REGULAR, INTERIOR = 0, 1, 2, 3


def render_svg(path, xy, tris, alpha, classes, size=1024, show_mesh=True, max_triangles=400000):
    """Write the picture to `path`.  xy [n, 2]; tris int32 [F, 9] (gpudtTriangle layout);
    alpha: the result dict (its edge order); classes: (vertex, edge, triangle) uint8 arrays."""
    xy = np.asarray(xy, dtype=np.float64)
    tris = np.asarray(tris).reshape(-1, 9)
    vcls, ecls, tcls = (np.asarray(c) for c in classes)
    if tris.shape[0] > max_triangles:
        raise ValueError(f"{tris.shape[0]} triangles: above max_triangles={max_triangles} for one SVG")
    lo, hi = xy.min(0), xy.max(0)
    span = float(max(hi[0] - lo[0], hi[1] - lo[1])) or 1.0
    margin = 8.0
    scale = (size - 2 * margin) / span

    def X(p):   # SVG y grows downwards
        return margin + (p[..., 0] - lo[0]) * scale, size - margin - (p[..., 1] - lo[1]) * scale

    out = [f'<svg xmlns="http://www.w3.org/2000/svg" width="{size}" height="{size}" '
           f'viewBox="0 0 {size} {size}">', '<rect width="100%" height="100%" fill="white"/>']
    V = tris[:, 3:6]
    tx, ty = X(xy[V])
    for t in np.nonzero(tcls == INTERIOR)[0]:
        pts = " ".join(f"{tx[t, k]:.2f},{ty[t, k]:.2f}" for k in range(3))
        out.append(f'<polygon class="interior" points="{pts}" fill="#9cc3e6" stroke="none"/>')
    et, ev = np.asarray(alpha["edge_t"]), np.asarray(alpha["edge_vopp"])
    a = tris[et, 3 + (ev + 1) % 3]
    b = tris[et, 3 + (ev + 2) % 3]
    ax, ay = X(xy[a])
    bx, by = X(xy[b])
    if show_mesh:
        for e in np.nonzero(ecls == EXTERIOR)[0]:
            out.append(f'<line class="mesh" x1="{ax[e]:.2f}" y1="{ay[e]:.2f}" x2="{bx[e]:.2f}" y2="{by[e]:.2f}" '
                       f'stroke="#dddddd" stroke-width="0.5"/>')
    for e in np.nonzero(ecls == SINGULAR)[0]:
        out.append(f'<line class="singular" x1="{ax[e]:.2f}" y1="{ay[e]:.2f}" x2="{bx[e]:.2f}" y2="{by[e]:.2f}" '
                   f'stroke="#d62728" stroke-width="1"/>')
    for e in np.nonzero(ecls == REGULAR)[0]:
        out.append(f'<line class="regular" x1="{ax[e]:.2f}" y1="{ay[e]:.2f}" x2="{bx[e]:.2f}" y2="{by[e]:.2f}" '
                   f'stroke="#1f3b73" stroke-width="2"/>')
    vx, vy = X(xy)
    for v in np.nonzero(vcls == SINGULAR)[0]:
        out.append(f'<circle class="singular" cx="{vx[v]:.2f}" cy="{vy[v]:.2f}" r="2" fill="#d62728"/>')
    out.append("</svg>")
    with open(path, "w") as f:
        f.write("\n".join(out))
