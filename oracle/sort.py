"""The reference's QuickSort (src/util/general.ih:6-121), oracle copy.

First-element pivot, Hoare-style hole filling, not stable: its tie order
fixes the quantum-number table order (real_tensor_space.cc:716-736) and the
truncation tie breaks (real_tensor_lattice.cc:538), so the exact partition
sequence matters.  flag=0 sorts descending, flag=1 ascending.  Written with
an explicit stack (the partition results do not depend on the visiting order
of independent sub-ranges), so deep recursion on presorted input is safe.
"""


def quicksort(arr, index=None, flag=0):
    """In-place sort of list ``arr`` (and the parallel list ``index``)."""
    n = len(arr)
    stack = [(0, n - 1)]
    while stack:
        left, right = stack.pop()
        if left >= right:
            continue
        i, j = left, right
        pivot = arr[i]
        pidx = index[i] if index is not None else None
        if flag == 0:
            while i < j:
                while i < j and arr[j] <= pivot:
                    j -= 1
                if i < j:
                    arr[i] = arr[j]
                    if index is not None:
                        index[i] = index[j]
                    i += 1
                while i < j and arr[i] >= pivot:
                    i += 1
                if i < j:
                    arr[j] = arr[i]
                    if index is not None:
                        index[j] = index[i]
                    j -= 1
        else:
            while i < j:
                while i < j and arr[j] >= pivot:
                    j -= 1
                if i < j:
                    arr[i] = arr[j]
                    if index is not None:
                        index[i] = index[j]
                    i += 1
                while i < j and arr[i] <= pivot:
                    i += 1
                if i < j:
                    arr[j] = arr[i]
                    if index is not None:
                        index[j] = index[i]
                    j -= 1
        arr[i] = pivot
        if index is not None:
            index[i] = pidx
        stack.append((i + 1, right))
        stack.append((left, i - 1))
    return arr, index
