// io.cpp -- *.mat / *.txt matrix I/O (include/io.hpp).  POSIX mmap for reading, a growable mmap append
// writer for output, hipHostMalloc-pinned dispatch buffers.
#include "io.hpp"
#include "mapped_file.hpp"

#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <bit>
#include <cerrno>
#include <cctype>
#include <cstring>
#include <iomanip>
#include <limits>
#include <locale>
#include <sstream>
#include <stdexcept>
#include <string>
#include <system_error>

namespace jacobi::svd::io
{
    namespace
    {
        constexpr std::size_t kHeaderBytes = 2 * sizeof(std::uint64_t);
        constexpr std::size_t kValueBytes = sizeof(std::uint64_t);
        constexpr std::size_t kMinPinnedCapacity = 64 * 1024;

        using jsvd::mapped::add_checked;
        using jsvd::mapped::be64;
        using jsvd::mapped::MappedAppendOutput;
        using jsvd::mapped::MappedInput;
        using jsvd::mapped::mul_checked;

        inline double load_be_double(const std::byte *p) noexcept
        {
            std::uint64_t raw;
            std::memcpy(&raw, p, sizeof(raw));
            return std::bit_cast<double>(be64(raw));
        }

        inline std::uint64_t load_be_u64(const std::byte *p) noexcept
        {
            std::uint64_t raw;
            std::memcpy(&raw, p, sizeof(raw));
            return be64(raw);
        }

        void check_layout(const Matrix &m)
        {
            if (m.values.size() != mul_checked(m.rows, m.columns))
            {
                throw std::invalid_argument("Matrix payload size does not match rows * columns.");
            }
        }

        std::size_t to_size(std::uint64_t v)
        {
            if (v > static_cast<std::uint64_t>(std::numeric_limits<std::size_t>::max()))
            {
                throw std::overflow_error("Matrix dimensions exceed platform size_t range.");
            }
            return static_cast<std::size_t>(v);
        }

        bool blank(const std::string &line)
        {
            return std::all_of(line.begin(), line.end(), [](unsigned char c) { return std::isspace(c) != 0; });
        }

        std::vector<double> parse_row(const std::string &line)
        {
            std::istringstream in(line);
            in.imbue(std::locale::classic());
            std::vector<double> row;
            double x = 0.0;
            while (in >> x)
            {
                row.push_back(x);
            }
            if (in.fail() && !in.eof())
            {
                throw std::invalid_argument("Invalid numeric token in text matrix row.");
            }
            return row;
        }
    } // namespace

    struct MatReader::Impl
    {
        MappedInput file;
        std::size_t cursor = 0;
        explicit Impl(const std::filesystem::path &path) : file(path) {}
    };

    struct MatWriter::Impl
    {
        MappedAppendOutput out;
        explicit Impl(const std::filesystem::path &path) : out(path) {}
    };

    struct TxtReader::Impl
    {
        std::ifstream in;
        explicit Impl(const std::filesystem::path &path) : in(path, std::ios::in | std::ios::binary)
        {
            if (!in)
            {
                throw std::runtime_error("Failed to open text matrix file for reading.");
            }
            in.imbue(std::locale::classic());
        }
    };

    struct TxtWriter::Impl
    {
        std::ofstream out;
        bool wrote_any = false;
        explicit Impl(const std::filesystem::path &path) : out(path, std::ios::out | std::ios::binary | std::ios::trunc)
        {
            if (!out)
            {
                throw std::runtime_error("Failed to open text matrix file for writing.");
            }
            out.imbue(std::locale::classic());
            out << std::setprecision(std::numeric_limits<double>::max_digits10);
        }
    };

    MatReader::MatReader(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    MatReader::~MatReader() = default;
    MatReader::MatReader(MatReader &&) noexcept = default;
    MatReader &MatReader::operator=(MatReader &&) noexcept = default;
    MatWriter::MatWriter(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    MatWriter::~MatWriter() = default;
    MatWriter::MatWriter(MatWriter &&) noexcept = default;
    MatWriter &MatWriter::operator=(MatWriter &&) noexcept = default;
    TxtReader::TxtReader(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    TxtReader::~TxtReader() = default;
    TxtReader::TxtReader(TxtReader &&) noexcept = default;
    TxtReader &TxtReader::operator=(TxtReader &&) noexcept = default;
    TxtWriter::TxtWriter(const std::filesystem::path &path) : impl_(std::make_unique<Impl>(path)) {}
    TxtWriter::~TxtWriter() = default;
    TxtWriter::TxtWriter(TxtWriter &&) noexcept = default;
    TxtWriter &TxtWriter::operator=(TxtWriter &&) noexcept = default;

    // ------------------------------------------------------------------------------ pinned task buffers

    PinnedHostTaskBuffer::~PinnedHostTaskBuffer() { release(); }

    PinnedHostTaskBuffer::PinnedHostTaskBuffer(PinnedHostTaskBuffer &&other) noexcept { move_from(std::move(other)); }

    PinnedHostTaskBuffer &PinnedHostTaskBuffer::operator=(PinnedHostTaskBuffer &&other) noexcept
    {
        if (this != &other)
        {
            release();
            move_from(std::move(other));
        }
        return *this;
    }

    void PinnedHostTaskBuffer::reserve(std::size_t input_bytes, std::size_t workspace_bytes)
    {
        const std::size_t need = add_checked(input_bytes, workspace_bytes);
        if (need > capacity_bytes_)
        {
            std::size_t cap = capacity_bytes_ == 0 ? kMinPinnedCapacity : capacity_bytes_;
            while (cap < need)
            {
                cap = (cap > std::numeric_limits<std::size_t>::max() / 2) ? need : cap * 2;
            }
            void *p = nullptr;
            const hipError_t st = hipHostMalloc(&p, cap, hipHostMallocDefault);
            if (st != hipSuccess)
            {
                throw std::runtime_error(std::string("hipHostMalloc failed: ") + hipGetErrorString(st));
            }
            release();
            data_ = static_cast<std::byte *>(p);
            capacity_bytes_ = cap;
        }
        input_size_bytes_ = input_bytes;
        workspace_size_bytes_ = workspace_bytes;
    }

    std::span<std::byte> PinnedHostTaskBuffer::mutable_input_bytes() noexcept { return {data_, input_size_bytes_}; }
    std::span<const std::byte> PinnedHostTaskBuffer::input_bytes() const noexcept { return {data_, input_size_bytes_}; }
    std::span<std::byte> PinnedHostTaskBuffer::mutable_workspace_bytes() noexcept
    {
        return {data_ + input_size_bytes_, workspace_size_bytes_};
    }
    std::span<const std::byte> PinnedHostTaskBuffer::workspace_bytes() const noexcept
    {
        return {data_ + input_size_bytes_, workspace_size_bytes_};
    }
    std::size_t PinnedHostTaskBuffer::capacity_bytes() const noexcept { return capacity_bytes_; }
    std::size_t PinnedHostTaskBuffer::input_size_bytes() const noexcept { return input_size_bytes_; }
    std::size_t PinnedHostTaskBuffer::workspace_size_bytes() const noexcept { return workspace_size_bytes_; }

    void PinnedHostTaskBuffer::release() noexcept
    {
        if (data_)
        {
            (void)hipHostFree(data_);
            data_ = nullptr;
        }
        capacity_bytes_ = input_size_bytes_ = workspace_size_bytes_ = 0;
    }

    void PinnedHostTaskBuffer::move_from(PinnedHostTaskBuffer &&other) noexcept
    {
        data_ = std::exchange(other.data_, nullptr);
        capacity_bytes_ = std::exchange(other.capacity_bytes_, 0);
        input_size_bytes_ = std::exchange(other.input_size_bytes_, 0);
        workspace_size_bytes_ = std::exchange(other.workspace_size_bytes_, 0);
    }

    // ------------------------------------------------------------------------------ dispatch reader

    MatDispatchReader::MatDispatchReader(const std::filesystem::path &path) : input_(path, std::ios::in | std::ios::binary)
    {
        if (!input_)
        {
            throw std::runtime_error("Failed to open *.mat file for dispatch reading.");
        }
    }

    bool MatDispatchReader::read_next(MatDispatchTask &task, std::size_t workspace_bytes)
    {
        std::byte header[kHeaderBytes];
        input_.read(reinterpret_cast<char *>(header), static_cast<std::streamsize>(kHeaderBytes));
        const std::streamsize got = input_.gcount();
        if (got == 0 && input_.eof())
        {
            return false;
        }
        if (got != static_cast<std::streamsize>(kHeaderBytes))
        {
            throw std::runtime_error("Truncated *.mat header while dispatch reading.");
        }
        const std::size_t rows = to_size(load_be_u64(header));
        const std::size_t cols = to_size(load_be_u64(header + kValueBytes));
        const std::size_t payload = mul_checked(mul_checked(rows, cols), kValueBytes);
        if (payload > static_cast<std::size_t>(std::numeric_limits<std::streamsize>::max()))
        {
            throw std::overflow_error("Byte size exceeds streamsize range.");
        }
        task.buffer.reserve(payload, workspace_bytes);
        if (payload > 0)
        {
            input_.read(reinterpret_cast<char *>(task.buffer.mutable_input_bytes().data()),
                        static_cast<std::streamsize>(payload));
            if (input_.gcount() != static_cast<std::streamsize>(payload))
            {
                throw std::runtime_error("Truncated *.mat payload while dispatch reading.");
            }
        }
        task.sequence_index = next_sequence_index_;
        task.rows = rows;
        task.columns = cols;
        next_sequence_index_ = add_checked(next_sequence_index_, 1);
        return true;
    }

    Matrix decode_dispatch_task_matrix(const MatDispatchTask &task)
    {
        Matrix m;
        m.rows = task.rows;
        m.columns = task.columns;
        const std::size_t count = mul_checked(m.rows, m.columns);
        const std::span<const std::byte> payload = task.buffer.input_bytes();
        if (payload.size() != mul_checked(count, kValueBytes))
        {
            throw std::invalid_argument("Dispatch payload size does not match rows * columns.");
        }
        m.values.resize(count);
        for (std::size_t i = 0; i < count; ++i)
        {
            m.values[i] = load_be_double(payload.data() + i * kValueBytes);
        }
        return m;
    }

    // ------------------------------------------------------------------------------ *.mat policy

    MatFilePolicy::Reader MatFilePolicy::open_reader(const std::filesystem::path &path) { return MatReader(path); }
    MatFilePolicy::Writer MatFilePolicy::open_writer(const std::filesystem::path &path) { return MatWriter(path); }

    bool MatFilePolicy::read_next(Reader &reader, Matrix &matrix)
    {
        if (!reader.impl_)
        {
            throw std::runtime_error("MatReader is not initialized.");
        }
        auto &st = *reader.impl_;
        const std::span<const std::byte> bytes = st.file.bytes();
        if (st.cursor == bytes.size())
        {
            return false;
        }
        if (bytes.size() - st.cursor < kHeaderBytes)
        {
            throw std::runtime_error("Truncated *.mat header.");
        }
        const std::size_t rows = to_size(load_be_u64(bytes.data() + st.cursor));
        const std::size_t cols = to_size(load_be_u64(bytes.data() + st.cursor + kValueBytes));
        st.cursor += kHeaderBytes;
        const std::size_t count = mul_checked(rows, cols);
        const std::size_t payload = mul_checked(count, kValueBytes);
        if (bytes.size() - st.cursor < payload)
        {
            throw std::runtime_error("Truncated *.mat payload.");
        }
        matrix.rows = rows;
        matrix.columns = cols;
        matrix.values.resize(count);
        const std::byte *src = bytes.data() + st.cursor;
        for (std::size_t i = 0; i < count; ++i)
        {
            matrix.values[i] = load_be_double(src + i * kValueBytes);
        }
        st.cursor += payload;
        return true;
    }

    void MatFilePolicy::write_next(Writer &writer, const Matrix &matrix)
    {
        if (!writer.impl_)
        {
            throw std::runtime_error("MatWriter is not initialized.");
        }
        check_layout(matrix);
        const std::uint64_t header[2] = {be64(static_cast<std::uint64_t>(matrix.rows)),
                                         be64(static_cast<std::uint64_t>(matrix.columns))};
        writer.impl_->out.append(header, sizeof(header));
        constexpr std::size_t kChunk = 32 * 1024;
        std::vector<std::uint64_t> chunk;
        chunk.reserve(kChunk);
        for (std::size_t done = 0; done < matrix.values.size();)
        {
            const std::size_t k = std::min(kChunk, matrix.values.size() - done);
            chunk.resize(k);
            for (std::size_t i = 0; i < k; ++i)
            {
                chunk[i] = be64(std::bit_cast<std::uint64_t>(matrix.values[done + i]));
            }
            writer.impl_->out.append(chunk.data(), k * sizeof(std::uint64_t));
            done += k;
        }
    }

    void MatFilePolicy::flush(Writer &writer)
    {
        if (!writer.impl_)
        {
            throw std::runtime_error("MatWriter is not initialized.");
        }
        writer.impl_->out.flush();
    }

    std::vector<Matrix> MatFilePolicy::read(const std::filesystem::path &path)
    {
        MatInputStream stream(path);
        return stream.read_all();
    }

    void MatFilePolicy::write(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        MatOutputStream stream(path);
        stream.write_all(matrices);
    }

    // ------------------------------------------------------------------------------ *.txt policy

    TxtFilePolicy::Reader TxtFilePolicy::open_reader(const std::filesystem::path &path) { return TxtReader(path); }
    TxtFilePolicy::Writer TxtFilePolicy::open_writer(const std::filesystem::path &path) { return TxtWriter(path); }

    bool TxtFilePolicy::read_next(Reader &reader, Matrix &matrix)
    {
        if (!reader.impl_)
        {
            throw std::runtime_error("TxtReader is not initialized.");
        }
        auto &in = reader.impl_->in;
        std::vector<double> values;
        std::size_t rows = 0;
        std::size_t cols = 0;
        bool inside = false;
        bool closed_by_blank = false;
        std::string line;
        while (std::getline(in, line))
        {
            if (blank(line))
            {
                if (!inside)
                {
                    continue;
                }
                closed_by_blank = true;
                break;
            }
            std::vector<double> row = parse_row(line);
            if (row.empty())
            {
                throw std::invalid_argument("Text matrix row cannot be empty when line is non-blank.");
            }
            if (!inside)
            {
                inside = true;
                cols = row.size();
            }
            else if (row.size() != cols)
            {
                throw std::invalid_argument("Inconsistent column count in text matrix block.");
            }
            values.insert(values.end(), row.begin(), row.end());
            rows = add_checked(rows, 1);
        }
        if (!closed_by_blank && !in.eof())
        {
            throw std::runtime_error("Failed while reading text matrix file.");
        }
        if (!inside)
        {
            return false;
        }
        matrix.rows = rows;
        matrix.columns = cols;
        matrix.values = std::move(values);
        return true;
    }

    void TxtFilePolicy::write_next(Writer &writer, const Matrix &matrix)
    {
        if (!writer.impl_)
        {
            throw std::runtime_error("TxtWriter is not initialized.");
        }
        check_layout(matrix);
        auto &st = *writer.impl_;
        if (st.wrote_any)
        {
            st.out << '\n';
        }
        for (std::size_t r = 0; r < matrix.rows; ++r)
        {
            for (std::size_t c = 0; c < matrix.columns; ++c)
            {
                if (c > 0)
                {
                    st.out << ' ';
                }
                st.out << matrix.values[r * matrix.columns + c];
            }
            st.out << '\n';
        }
        if (!st.out)
        {
            throw std::runtime_error("Failed while writing text matrix file.");
        }
        st.wrote_any = true;
    }

    void TxtFilePolicy::flush(Writer &writer)
    {
        if (!writer.impl_)
        {
            throw std::runtime_error("TxtWriter is not initialized.");
        }
        writer.impl_->out.flush();
        if (!writer.impl_->out)
        {
            throw std::runtime_error("Failed to flush text matrix file.");
        }
    }

    std::vector<Matrix> TxtFilePolicy::read(const std::filesystem::path &path)
    {
        TxtInputStream stream(path);
        return stream.read_all();
    }

    void TxtFilePolicy::write(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        TxtOutputStream stream(path);
        stream.write_all(matrices);
    }

    std::vector<Matrix> read_mat_file(const std::filesystem::path &path) { return MatFilePolicy::read(path); }
    void write_mat_file(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        MatFilePolicy::write(path, matrices);
    }
    std::vector<Matrix> read_txt_file(const std::filesystem::path &path) { return TxtFilePolicy::read(path); }
    void write_txt_file(const std::filesystem::path &path, std::span<const Matrix> matrices)
    {
        TxtFilePolicy::write(path, matrices);
    }
} // namespace jacobi::svd::io
