"""Independent brute-force tree enumerator used to pin the oracle.

It generates every expression tree of a given size straight from the grammar
definition, in the reference's enumeration order (nocache.cu:23-83 / 117-160):
operators in the order ~ & | ^ neg + - * ; for binary operators the left child
size j runs upward, with j <= (size-1)/2 for the commutative operators and
all j for '-'; inside one (op, j) the left tree varies slowest.  It shares no
code and no arithmetic with the oracle's count-table/unrank implementation.
"""
NOT, AND, OR, XOR, NEG, ADD, SUB, MUL = range(8)
TOK_OP = 32


def trees(size, L):
    if size == 1:
        for v in range(L):
            yield (v,)
        return
    for op in range(8):
        if op in (NOT, NEG):
            for c in trees(size - 1, L):
                yield c + (TOK_OP + op,)
        else:
            last = size - 2 if op == SUB else (size - 1) // 2
            for j in range(1, last + 1):
                rights = list(trees(size - 1 - j, L))
                for a in trees(j, L):
                    for b in rights:
                        yield a + b + (TOK_OP + op,)


def py_eval(rpn, leafvals, W):
    m = (1 << W) - 1
    st = []
    for t in rpn:
        if t < TOK_OP:
            st.append(leafvals[t] & m)
            continue
        op = t - TOK_OP
        if op == NOT:
            st.append(~st.pop() & m)
        elif op == NEG:
            st.append(-st.pop() & m)
        else:
            r = st.pop()
            l = st.pop()
            st.append({AND: l & r, OR: l | r, XOR: l ^ r, ADD: l + r,
                       SUB: l - r, MUL: l * r}[op] & m)
    assert len(st) == 1
    return st[0]
