// mapped_file.hpp -- POSIX memory-mapped input files and growable memory-mapped append output, shared by the
// *.mat reader/writer (io.cpp) and the streaming *.mat solver (mat_stream.cpp).
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <bit>
#include <cerrno>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <limits>
#include <span>
#include <stdexcept>
#include <system_error>

namespace jsvd::mapped
{
    constexpr std::size_t kMinMappedCapacity = 1 << 20;

    inline std::size_t mul_checked(std::size_t a, std::size_t b)
    {
        if (a != 0 && b > std::numeric_limits<std::size_t>::max() / a)
        {
            throw std::overflow_error("Size multiplication overflow.");
        }
        return a * b;
    }

    inline std::size_t add_checked(std::size_t a, std::size_t b)
    {
        if (a > std::numeric_limits<std::size_t>::max() - b)
        {
            throw std::overflow_error("Size addition overflow.");
        }
        return a + b;
    }

    // Host <-> big-endian ("network order") conversion.
    inline std::uint64_t be64(std::uint64_t x) noexcept
    {
        if constexpr (std::endian::native == std::endian::little)
        {
            return __builtin_bswap64(x);
        }
        return x;
    }

    // Read-only mapping of a whole file (sequential access advised).
    class MappedInput
    {
    public:
        explicit MappedInput(const std::filesystem::path &path)
        {
            fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
            if (fd_ < 0)
            {
                throw std::system_error(errno, std::system_category(), "open failed: " + path.string());
            }
            struct stat st{};
            if (::fstat(fd_, &st) != 0)
            {
                const int err = errno;
                ::close(fd_);
                throw std::system_error(err, std::system_category(), "fstat failed");
            }
            size_ = static_cast<std::size_t>(st.st_size);
            if (size_ > 0)
            {
                void *p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
                if (p == MAP_FAILED)
                {
                    const int err = errno;
                    ::close(fd_);
                    throw std::system_error(err, std::system_category(), "mmap failed");
                }
                (void)::madvise(p, size_, MADV_SEQUENTIAL);
                data_ = static_cast<const std::byte *>(p);
            }
        }
        ~MappedInput()
        {
            if (data_)
            {
                ::munmap(const_cast<std::byte *>(data_), size_);
            }
            if (fd_ >= 0)
            {
                ::close(fd_);
            }
        }
        MappedInput(const MappedInput &) = delete;
        MappedInput &operator=(const MappedInput &) = delete;

        [[nodiscard]] std::span<const std::byte> bytes() const noexcept { return {data_, size_}; }

    private:
        int fd_ = -1;
        const std::byte *data_ = nullptr;
        std::size_t size_ = 0;
    };

    // Append-only writer over a file mapping that doubles its capacity; flush() msyncs the written prefix, the
    // destructor truncates the file to the bytes written.
    class MappedAppendOutput
    {
    public:
        explicit MappedAppendOutput(const std::filesystem::path &path)
        {
            fd_ = ::open(path.c_str(), O_RDWR | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
            if (fd_ < 0)
            {
                throw std::system_error(errno, std::system_category(), "open failed: " + path.string());
            }
        }
        ~MappedAppendOutput() { close_quietly(); }
        MappedAppendOutput(const MappedAppendOutput &) = delete;
        MappedAppendOutput &operator=(const MappedAppendOutput &) = delete;

        // Reserves `bytes` at the end of the file and returns where to write them.
        std::byte *append_region(std::size_t bytes)
        {
            const std::size_t need = add_checked(size_, bytes);
            if (need > capacity_)
            {
                std::size_t cap = std::max(capacity_, kMinMappedCapacity);
                while (cap < need)
                {
                    cap = (cap > std::numeric_limits<std::size_t>::max() / 2) ? need : cap * 2;
                }
                remap(cap);
            }
            std::byte *dst = data_ + size_;
            size_ = need;
            return dst;
        }

        // Sizes the mapping for `bytes` in total up front (a writer that knows its output size): later appends up
        // to that size neither remap nor flush the written prefix.  No effect when the capacity already suffices.
        void reserve(std::size_t bytes)
        {
            if (bytes > capacity_)
            {
                remap(bytes);
            }
        }

        void append(const void *src, std::size_t bytes)
        {
            if (bytes == 0)
            {
                return;
            }
            std::memcpy(append_region(bytes), src, bytes);
        }

        void flush()
        {
            if (data_ && size_ > 0 && ::msync(data_, size_, MS_SYNC) != 0)
            {
                throw std::system_error(errno, std::system_category(), "msync failed");
            }
        }

    private:
        void remap(std::size_t cap)
        {
            if (data_)
            {
                flush();
                ::munmap(data_, capacity_);
                data_ = nullptr;
            }
            if (::ftruncate(fd_, static_cast<off_t>(cap)) != 0)
            {
                throw std::system_error(errno, std::system_category(), "ftruncate failed");
            }
            void *p = ::mmap(nullptr, cap, PROT_READ | PROT_WRITE, MAP_SHARED, fd_, 0);
            if (p == MAP_FAILED)
            {
                throw std::system_error(errno, std::system_category(), "mmap failed");
            }
            data_ = static_cast<std::byte *>(p);
            capacity_ = cap;
        }

        void close_quietly() noexcept
        {
            try
            {
                flush();
            }
            catch (...)
            {
            }
            if (data_)
            {
                ::munmap(data_, capacity_);
                data_ = nullptr;
            }
            if (fd_ >= 0)
            {
                (void)::ftruncate(fd_, static_cast<off_t>(size_));
                ::close(fd_);
                fd_ = -1;
            }
        }

        int fd_ = -1;
        std::byte *data_ = nullptr;
        std::size_t size_ = 0;
        std::size_t capacity_ = 0;
    };
} // namespace jsvd::mapped
