"""Crafted ladder-rich positions: the move lists (passes give the colours)
used by the golden fixture and the encoder tests."""
import numpy as np


def crafted_moves(seed, side=19):
    """Move list (passes give the colours) of a ladder-rich position: 1-4
    ladder starters (a string of the side to move in atari with its escape
    open, or an opponent stone with two liberties), then random stones of
    both colours at density 3-20% as breakers and ladder makers. Moves are
    chosen with the library's board (empty, not suicide); the expected
    values come from replaying them on the reference engine."""
    from leela_amd.positions import Position
    rng = np.random.default_rng(seed)
    p = Position(side)
    moves = []

    def place(c, x, y):
        if p.squares()[y * side + x] != 2:
            return
        if p.info()["to_move"] != c:
            p.play(-1)
            moves.append(-1)
        if p.queries()[y * side + x][9] == 1:
            return
        v = p.vertex(x, y)
        p.play(v)
        moves.append(v)

    tm = int(rng.integers(0, 2))
    for _ in range(rng.integers(1, 5)):
        x, y = (int(t) for t in rng.integers(2, side - 2, 2))
        dx, dy = (int(t) for t in rng.choice([-1, 1], 2))
        d = tm if rng.random() < 0.5 else 1 - tm
        place(d, x, y)
        place(1 - d, x + dx, y)
        place(1 - d, x, y + dy)
        if d == tm:
            place(1 - d, x - dx, y)
    for _ in range(int(rng.uniform(0.03, 0.2) * side * side)):
        c = int(rng.integers(0, 2))
        x, y = (int(t) for t in rng.integers(0, side, 2))
        place(c, x, y)
    if p.info()["to_move"] != tm:
        p.play(-1)
        moves.append(-1)
    return np.array(moves, np.int32)
