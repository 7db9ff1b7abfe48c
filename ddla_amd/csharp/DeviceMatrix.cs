// Device-resident FP64 matrices for the GPU path of DDLA: a SafeHandle over the library's
// device allocation (released with ddla_free even when the owner forgets Dispose), a row-major
// matrix with a leading dimension, uploads and downloads from strided host views, and the
// factorizations and solves that keep their operands on the device between calls.
using System;
using System.Runtime.InteropServices;

namespace DDLA.Hip;

/// <summary>An error status of libddla_hip.so with its message.</summary>
public sealed class HipException : Exception
{
    public int Status { get; }
    public HipException(int status, string message) : base($"status {status}: {message}")
        => Status = status;
}

/// <summary>A factorization found a zero (or, for Cholesky, non-positive) pivot; the same
/// condition for which the reference throws LinalgException (Cholesky.cs:228-236) or
/// ArgumentOutOfRangeException from InvScal (BLAS/1v.cs:99).</summary>
public sealed class HipPivotException : Exception
{
    public long Index { get; }
    public HipPivotException(string what, long index)
        : base($"{what}: zero or non-positive pivot at {index}") => Index = index;
}

public static class HipStatus
{
    /// <summary>Throws on a negative status; returns the non-negative info.</summary>
    public static int Check(int status)
    {
        if (status < 0) throw new HipException(status, HipNative.LastError());
        return status;
    }
}

/// <summary>Owns one device allocation of the library (ddla_malloc / ddla_free).</summary>
public sealed class DeviceMemoryHandle : SafeHandle
{
    public DeviceMemoryHandle() : base(IntPtr.Zero, true) { }

    public static DeviceMemoryHandle Allocate(long bytes)
    {
        HipStatus.Check(HipNative.ddla_malloc(out IntPtr p, (nuint)Math.Max(bytes, 8)));
        var h = new DeviceMemoryHandle();
        h.SetHandle(p);
        return h;
    }

    public override bool IsInvalid => handle == IntPtr.Zero;

    protected override bool ReleaseHandle() => HipNative.ddla_free(handle) == 0;
}

/// <summary>A rows x cols row-major FP64 matrix in device memory, element (i, j) at
/// Ptr[i * Ld + j].</summary>
public sealed class DeviceMatrix : IDisposable
{
    private readonly DeviceMemoryHandle memory;
    public long Rows { get; }
    public long Cols { get; }
    public long Ld { get; }
    public IntPtr Ptr => memory.DangerousGetHandle();

    public DeviceMatrix(long rows, long cols)
    {
        if (rows < 0 || cols < 0) throw new ArgumentOutOfRangeException(nameof(rows));
        Rows = rows;
        Cols = cols;
        Ld = Math.Max(cols, 1);
        memory = DeviceMemoryHandle.Allocate(rows * Ld * sizeof(double));
    }

    /// <summary>Copies the strided host view (rowStride, colStride in elements) starting at
    /// head into a new device matrix.</summary>
    public static DeviceMatrix Upload(long rows, long cols, ref double head, long rowStride,
                                      long colStride)
    {
        var d = new DeviceMatrix(rows, cols);
        unsafe
        {
            fixed (double* p = &head)
                HipStatus.Check(HipNative.ddla_dmat_set(rows, cols, (IntPtr)p, rowStride,
                                                        colStride, d.Ptr, d.Ld));
        }
        return d;
    }

    public static DeviceMatrix Upload(double[,] a)
    {
        long rows = a.GetLength(0), cols = a.GetLength(1);
        if (a.Length == 0) return new DeviceMatrix(rows, cols);
        return Upload(rows, cols, ref a[0, 0], cols, 1);
    }

    /// <summary>Copies the device matrix into the strided host view at head.</summary>
    public void Download(ref double head, long rowStride, long colStride)
    {
        unsafe
        {
            fixed (double* p = &head)
                HipStatus.Check(HipNative.ddla_dmat_get(Rows, Cols, Ptr, Ld, (IntPtr)p, rowStride,
                                                        colStride));
        }
    }

    public double[,] ToArray()
    {
        var a = new double[Rows, Cols];
        if (a.Length > 0) Download(ref a[0, 0], Cols, 1);
        return a;
    }

    public void Dispose() => memory.Dispose();
}

/// <summary>Pivots (int32) in device memory, as the device factorizations return them.</summary>
public sealed class DevicePivots : IDisposable
{
    private readonly DeviceMemoryHandle memory;
    public long Length { get; }
    public IntPtr Ptr => memory.DangerousGetHandle();

    public DevicePivots(long n)
    {
        Length = n;
        memory = DeviceMemoryHandle.Allocate(Math.Max(n, 1) * sizeof(int));
    }

    public int[] ToArray()
    {
        var p = new int[Length];
        unsafe
        {
            fixed (int* q = p)
                HipStatus.Check(HipNative.ddla_memcpy_d2h((IntPtr)q, Ptr, (nuint)(Length * sizeof(int))));
        }
        return p;
    }

    public void Dispose() => memory.Dispose();
}
