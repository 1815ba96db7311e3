"""Writes tests/golden/examples_fast.npz from the reference's own fixture data.

Run in the build container (reads /root/reference); the .npz travels.
Sources:
  examples/examples_fast/Data/xpsi_good_realisation.dat   (counts, 291 x 32)
  examples/examples_fast/Data/xpsi_synthetic_background.dat (background)
  xpsi/tests/test_likelihood.py:34-65 and
  examples/examples_fast/Modules/main.py:33-82 (instrument recipe: ARF edges
  accumulated by repeated float addition, diagonal RMF, ARF 1800 cm^2;
  CustomInstrument.from_response_files, CustomInstrument.py:29-62).
"""
import os
import numpy as np

REF = "/root/reference/examples/examples_fast/Data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples_fast.npz")


def instrument():
    # ARF bin edges are built by repeated addition of 0.005 keV, exactly as
    # the reference recipe does; linspace would give different last bits.
    lo, hi, val = [0.1], [0.105], [1800]
    counter = 1
    while lo[-1] <= 14.995:
        lo.append(lo[-1] + 0.005)
        hi.append(hi[-1] + 0.005)
        val.append(1800)
        counter += 1
    ARF = np.array([lo, hi, val], dtype=np.float64).T
    RMF = np.diag(np.full(counter, 1))
    min_input, max_input, ch = 10, 301, (10, 301)
    matrix = np.ascontiguousarray(RMF[min_input:max_input, ch[0]:ch[1]].T,
                                  dtype=np.double)
    edges = np.zeros(ARF[min_input:max_input, 2].shape[0] + 1)
    edges[0] = ARF[min_input, 0]
    edges[1:] = ARF[min_input:max_input, 1]
    for i in range(matrix.shape[0]):
        matrix[i, :] *= ARF[min_input:max_input, 2]
    return matrix, edges


def main():
    counts = np.loadtxt(os.path.join(REF, "xpsi_good_realisation.dat"), dtype=np.double)
    bkg = np.loadtxt(os.path.join(REF, "xpsi_synthetic_background.dat"))
    matrix, edges = instrument()
    np.savez_compressed(OUT, counts=counts, background_counts=bkg[:, 1],
                        matrix_diag=np.diag(matrix).copy(), energy_edges=edges)
    print("wrote", OUT, counts.shape, edges.shape)


if __name__ == "__main__":
    main()
