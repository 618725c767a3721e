"""Seeded FILL recipes for the initialization tests: every method of
filling_phase_at (pearl_serial.cpp:1393-1514), with parameters inside
read_config's validation rules (validate_filling_operations :1251).

A recipe opens with one whole-map operation (centre box or a Y junction,
which set every cell) and continues with overlays (cubes, discs, pearlite
arms, ellipses) that change only their masks, so every method leaves cells
of its own in the final labels."""
import random

WHOLE_MAP = ("FILLCENTERBOX2D", "FILLYJUNCTION2D", "FILLYJUNCTIONLAMELLAE2D")


def _text(vals):
    return "{" + ",".join(repr(float(v)) if isinstance(v, float) else str(int(v)) for v in vals) + "}"


def _box(r, nx, ny):
    x0 = r.randrange(0, nx // 2)
    y0 = r.randrange(0, ny // 2)
    return x0, r.randrange(x0 + nx // 4, nx + 1), y0, r.randrange(y0 + ny // 4, ny + 1)


def recipe(nx: int, ny: int, seed: int, whole_map: str, overlays: int = 6):
    """[(method, "{...}")] for inputs.write_input."""
    r = random.Random(seed)
    out = []
    if whole_map == "FILLCENTERBOX2D":
        a, b = r.sample([0, 1, 2], 2)
        out.append((whole_map, [*_box(r, nx, ny), a, b]))
    else:
        lam = whole_map == "FILLYJUNCTIONLAMELLAE2D"
        box = _box(r, nx, ny)
        cx, cy = r.randrange(box[0], box[1]), r.randrange(box[2], box[3])
        base = r.uniform(0.0, 120.0)
        angles = [base + 120.0 * k + r.uniform(-15.0, 15.0) for k in range(3)]
        length, half = r.uniform(0.3, 0.8) * max(nx, ny), r.uniform(2.0, 0.15 * min(nx, ny))
        if lam:
            f, c, g = r.sample([0, 1, 2], 3)
            out.append((whole_map, [*box, cx, cy, *angles, length, half, r.uniform(3.0, 9.0), r.uniform(0.8, 3.0),
                                    f, c, g]))
        else:
            f, g = r.sample([0, 1, 2], 2)
            out.append((whole_map, [*box, cx, cy, *angles, length, half, f, g]))
    for k in range(overlays):
        kind = k % 4
        if kind == 0:  # cube, gamma included (a no-op, :1400-1401)
            x0, y0 = r.randrange(0, nx), r.randrange(0, ny)
            out.append(("FILLCUBE", [r.choice([0, 1, 2]), x0, y0, 0, r.randrange(x0, nx), r.randrange(y0, ny), 0]))
        elif kind == 1:  # disc
            out.append(("FILLCYLINDER", [r.choice([0, 1, 2]), r.randrange(0, nx), r.randrange(0, ny), 0, 0,
                                         r.uniform(1.0, 0.3 * min(nx, ny))]))
        elif kind == 2:  # adhered pearlite arm
            f, c = r.sample([0, 1, 2], 2)
            peak = r.uniform(2.0, 0.2 * min(nx, ny))
            out.append(("FILLADHEREDPEARLITEARM2D",
                        [r.uniform(0.0, nx - 1.0), r.uniform(0.0, ny - 1.0), r.uniform(-180.0, 180.0),
                         r.uniform(0.3, 0.9) * max(nx, ny), r.randint(1, 7), r.uniform(0.1, 0.4), peak,
                         r.uniform(0.0, peak), r.choice([0.0, r.uniform(1.0, 6.0)]), f, c]))
        else:  # pearlite ellipse
            f, c = r.sample([0, 1, 2], 2)
            out.append(("FILLPEARLITEELLIPSE2D",
                        [r.uniform(0.0, nx - 1.0), r.uniform(0.0, ny - 1.0), r.uniform(-180.0, 180.0),
                         r.uniform(0.0, 10.0), r.randint(1, 6), r.uniform(4.0, 12.0), r.uniform(0.2, 0.8),
                         r.uniform(0.1, 0.4), r.choice([0.0, r.uniform(1.0, 5.0)]), f, c]))
    return [(m, _text(v)) for m, v in out]
