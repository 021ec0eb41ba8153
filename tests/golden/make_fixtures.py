"""Write ``fixtures.json``: known answers held by the reference's own files.

Values are transcribed from notebook outputs and test constants of
DaniloZZZ/QTensor (paths relative to the reference checkout):

* F1  examples/QTensor usage.ipynb:164-181,295-309 -- P4 graph with edges
      (0,3),(1,2),(2,3); QtreeQAOAComposer.ansatz_state, p=1, gamma=.1, beta=.2;
      amplitude <0000|U|0000> = 0.08550212+0.38739031j (Cirq: 0.085502125+0.38739026j, :352).
* F2  examples/QTensor usage.ipynb:487-508 -- same graph, p=2, gamma=[.1,.1],
      beta=[.2,.2]; QAOAQtreeSimulator.energy_expectation = 1.52511548.
* F3  examples/Qtree Basic usage.ipynb:186,296,368,476,755 -- one-Hadamard network:
      bucket layout, full contraction sqrt(2), ket-sliced sqrt(2), <0|H|0> = 0.7071067811865475,
      partial contraction over the output wire [0.70710678, 0.70710678].
* F4  qtensor/tests/qiskit_qaoa_energy.py:112-123 -- 10-node, 15-edge graph, p=9
      Qiskit angles; expected cut 12 within 1e-2 (shot-based).  Qiskit's rzz(t) and
      rx(2 b) map to QTensor units gamma = t/(2 pi), beta = b/pi.

Run ``python tests/golden/make_fixtures.py`` to regenerate.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

FIXTURES = {
    "F1": {"edges": [[0, 3], [1, 2], [2, 3]], "n": 4, "gamma": [0.1], "beta": [0.2],
           "amplitude": [0.08550212, 0.38739031], "amplitude_cirq": [0.085502125, 0.38739026],
           "digits": 8},
    "F2": {"edges": [[0, 3], [1, 2], [2, 3]], "gamma": [0.1, 0.1], "beta": [0.2, 0.2],
           "energy": 1.52511548, "digits": 8},
    "F3": {"buckets": "[[M(o_0,v_1)], [MyGate(v_1,v_2)], [M(v_2,i_0)], []]",
           "full": 1.414213562373095, "ket_sliced": 1.414213562373095,
           "amplitude": 0.7071067811865475, "partial": [0.70710678, 0.70710678]},
    "F4": {"edges": [[0, 1], [1, 2], [2, 3], [3, 4], [4, 0], [0, 5], [1, 6], [2, 7], [3, 8],
                     [4, 9], [5, 7], [5, 8], [6, 8], [6, 9], [7, 9]],
           "qiskit_beta": [5.192253984583296, 5.144373231492732, 5.9438949617723775,
                           5.807748946652058, 3.533458907810596, 6.006206583282401,
                           6.122313961527631, 6.218468942101044, 6.227704753217614],
           "qiskit_rzz": [0.3895570099244132, -0.1809282325810937, 0.8844522327007089,
                          0.7916086532373585, 0.21294534589417236, 0.4328896243354414,
                          0.8327451563500539, 0.7694639329585451, 0.4727893829336214],
           "cut": 12.0, "tol": 1e-2},
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(FIXTURES, f, indent=1)
    print("wrote", os.path.join(HERE, "fixtures.json"))
