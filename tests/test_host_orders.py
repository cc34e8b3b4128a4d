"""Index arithmetic of the main-force latency kernel's non-point pass
(k_main_wpb, eph_kernels.hip), checked against the loop structure of the
pooled non-point rounds it reproduces (src/physics/mass.cpp:283-297): round
k of body i's chunk sj visits item chunk (sj - k) mod T, each chunk from its
last item down, where chunk c holds items [c*nit/T, (c+1)*nit/T) and body i
lies in partner chunk sj, the chunk c with c*n/T <= i < (c+1)*n/T."""


def chunk_of(i, n, T):
    sj = 0
    while (sj + 1) * n // T <= i:
        sj += 1
    return sj


def test_partner_chunk_closed_form():
    # the kernel computes sj = ((i + 1) * T - 1) / n
    for n in range(1, 161):
        for T in range(1, min(n, 64) + 1):
            for i in range(n):
                assert ((i + 1) * T - 1) // n == chunk_of(i, n, T), (n, T, i)


def test_receiver_order_is_one_cyclic_descending_pass():
    # the kernel visits q = (sj+1)*nit/T - 1, q-1, ... (mod nit), nit items
    for nit in range(1, 65):
        for T in range(1, nit + 1):          # T = min(items, tasks) <= nit
            for sj in range(T):
                rounds = []
                for k in range(T):
                    cc = (sj - k) % T
                    first, last = cc * nit // T, (cc + 1) * nit // T
                    rounds += list(range(last - 1, first - 1, -1))
                q0 = (sj + 1) * nit // T - 1
                assert rounds == [(q0 - m) % nit for m in range(nit)], (nit, T, sj)
